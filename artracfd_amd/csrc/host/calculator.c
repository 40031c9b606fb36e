/*
 * calculator.c - initial-condition expressions.
 *
 * Expressions follow the solver's calculator grammar (calculator.c:126-718
 * of the reference): numbers written as digits[.digits], variables t x y z
 * pi ans, binary + - * / ^, functions exp ln lg abs sin cos tan, and three
 * bracket kinds () [] {}.  A + or - right after an opening bracket (or at
 * the start) is unary and acts as 0 +/- operand.  Precedence, lowest first:
 * {+ -} < {* /} < {^} < functions; every binary level is left-associative
 * (2^3^2 is 64) and a function binds tighter than ^ (sin x^2 is (sin x)^2).
 * This reproduces the reference's operator-priority matrix
 * (calculator.c:139-158), including its error cases.
 *
 * The reference re-parses the text for every grid node.  Here the text is
 * compiled once into a postfix program whose evaluation performs the same
 * floating-point operations in the same order, so results are identical.
 * As in the reference, any error leaves the answer untouched.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "artra_host.h"

enum { MAXSTACK = 49, MAXPROG = 256, MAXTEXT = STR };

typedef enum { OP_NUM, OP_VAR, OP_UNARY, OP_BINARY, OP_FAIL } OpKind;

typedef struct {
    OpKind kind;
    char sym;      /* operator symbol or variable letter after translation */
    Real value;    /* OP_NUM */
} Op;

struct ArtraExpr {
    int n;
    int broken;    /* the text itself is rejected: evaluation always fails */
    Op op[MAXPROG];
};

static int IsDigit(char c) { return c >= '0' && c <= '9'; }
static int IsLeft(char c) { return c == '(' || c == '[' || c == '{'; }
static int IsRight(char c) { return c == ')' || c == ']' || c == '}'; }
static int IsFunc(char c) { return c == 'e' || c == 'n' || c == 'g' || c == 'a' || c == 's' || c == 'c' || c == 't'; }
static int IsBinary(char c) { return c == '+' || c == '-' || c == '*' || c == '/' || c == '^'; }
static int IsVar(char c) { return c == 'u' || c == 'x' || c == 'y' || c == 'z' || c == 'p' || c == 'q'; }

/* Rewrite names to one-letter symbols (calculator.c:405-538). */
static int Translate(char *s)
{
    char *r = s, *w = s;
    while (*r) {
        switch (*r) {
            case 'e':
                if ('x' == r[1] && 'p' == r[2]) { *w++ = 'e'; r += 3; } else return 1;
                break;
            case 'l':
                if ('n' == r[1]) { *w++ = 'n'; r += 2; }
                else if ('g' == r[1]) { *w++ = 'g'; r += 2; }
                else return 1;
                break;
            case 'a':
                if ('b' == r[1] && 's' == r[2]) { *w++ = 'a'; r += 3; }
                else if ('n' == r[1] && 's' == r[2]) { *w++ = 'q'; r += 3; }
                else return 1;
                break;
            case 's':
                if ('i' == r[1] && 'n' == r[2]) { *w++ = 's'; r += 3; } else return 1;
                break;
            case 'c':
                if ('o' == r[1] && 's' == r[2]) { *w++ = 'c'; r += 3; } else return 1;
                break;
            case 't':
                if ('a' == r[1] && 'n' == r[2]) { *w++ = 't'; r += 3; } else { *w++ = 'u'; r += 1; }
                break;
            case 'p':
                if ('i' == r[1]) { *w++ = 'p'; r += 2; } else return 1;
                break;
            case 'x': case 'y': case 'z':
                *w++ = *r++;
                break;
            case '.':
                /* r[-1] reads the character before the text at the start: the
                 * caller places an opening bracket there */
                if (IsDigit(r[1]) && IsDigit(r[-1])) { *w++ = *r++; } else return 1;
                break;
            case ' ':
                ++r;
                break;
            default:
                if (IsDigit(*r) || IsBinary(*r) || IsLeft(*r) || IsRight(*r)) { *w++ = *r++; } else return 1;
                break;
        }
    }
    *w = '\0';
    return 0;
}

/* precedence level of an operator on the stack / incoming */
static int Level(char c)
{
    if ('+' == c || '-' == c) return 1;
    if ('*' == c || '/' == c) return 2;
    if ('^' == c) return 3;
    return 4; /* functions */
}

/* '<' shift, '>' reduce, '=' drop a matched bracket pair, 'w' error */
static char Relation(char top, char in)
{
    if (IsLeft(top) || '\0' == top) {
        if (IsRight(in) || '\0' == in) {
            if (('(' == top && ')' == in) || ('[' == top && ']' == in) || ('{' == top && '}' == in) ||
                ('\0' == top && '\0' == in)) {
                return '=';
            }
            return 'w';
        }
        return '<';
    }
    /* top is a binary operator or a function */
    if (IsRight(in) || '\0' == in) return '>';
    if (IsLeft(in) || IsFunc(in)) return '<';
    return (Level(top) >= Level(in)) ? '>' : '<';
}

static int Emit(ArtraExpr *e, OpKind kind, char sym, Real value, int *depth)
{
    if (e->n >= MAXPROG) return 1;
    e->op[e->n].kind = kind;
    e->op[e->n].sym = sym;
    e->op[e->n].value = value;
    e->n++;
    switch (kind) {
        case OP_NUM: case OP_VAR:
            if (*depth >= MAXSTACK) return 1; /* operand stack overflow */
            ++*depth;
            break;
        case OP_UNARY:
            if (*depth < 1) return 1;
            break;
        case OP_BINARY:
            if (*depth < 2) return 1;
            --*depth;
            break;
        default:
            break;
    }
    return 0;
}

static int Compile(const char *text, ArtraExpr *e)
{
    char buf[MAXTEXT + 8];
    buf[0] = '(';
    strncpy(buf + 1, text, MAXTEXT - 5);
    buf[MAXTEXT - 4] = '\0';
    char *s = buf + 1;
    if (Translate(s)) return 1;
    char ops[MAXSTACK + 1];
    int top = 0;
    int depth = 0;
    ops[top++] = '\0';
    while ('\0' != *s || '\0' != ops[top - 1]) {
        if (IsDigit(*s)) {
            char *end = s;
            while (IsDigit(*end)) ++end;
            if ('.' == *end) ++end;
            while (IsDigit(*end)) ++end;
            char num[MAXTEXT];
            memcpy(num, s, end - s);
            num[end - s] = '\0';
            if (Emit(e, OP_NUM, 0, strtod(num, NULL), &depth)) return 1;
            s = end;
            continue;
        }
        if (IsVar(*s)) {
            if (Emit(e, OP_VAR, *s, 0.0, &depth)) return 1;
            ++s;
            continue;
        }
        const char in = *s;
        if (!(IsFunc(in) || IsBinary(in) || IsLeft(in) || IsRight(in) || '\0' == in)) return 1;
        switch (Relation(ops[top - 1], in)) {
            case '<':
                if (('+' == in || '-' == in) && IsLeft(s[-1])) {
                    if (Emit(e, OP_NUM, 0, 0.0, &depth)) return 1;
                }
                if (top >= MAXSTACK) return 1; /* operator stack overflow */
                ops[top++] = in;
                ++s;
                break;
            case '=':
                --top;
                ++s;
                break;
            case '>': {
                const char op = ops[--top];
                if (Emit(e, IsFunc(op) ? OP_UNARY : OP_BINARY, op, 0.0, &depth)) return 1;
                break;
            }
            default:
                return 1; /* brackets do not match */
        }
    }
    return (1 == depth) ? 0 : 1;
}

int artra_expr_compile(const char *text, ArtraExpr **out)
{
    ArtraExpr *e = calloc(1, sizeof(*e));
    if (NULL == e) return artra_set_error(1, "out of host memory");
    if (Compile(text, e)) {
        e->broken = 1;
        e->n = 0;
    }
    *out = e;
    return 0;
}

void artra_expr_free(ArtraExpr *e)
{
    free(e);
}

/* Unary and binary operations with the reference's domain checks (calculator.c:642-718). */
static int Unary(char op, Real x, Real *r)
{
    switch (op) {
        case 'e': *r = exp(x); return 0;
        case 'n': if (0.0 >= x) return 1; *r = log(x); return 0;
        case 'g': if (0.0 >= x) return 1; *r = log10(x); return 0;
        case 'a': *r = fabs(x); return 0;
        case 's': *r = sin(x); return 0;
        case 'c': *r = cos(x); return 0;
        case 't': if (0.0 == cos(x)) return 1; *r = sin(x) / cos(x); return 0;
        default: return 1;
    }
}

static int Binary(Real a, char op, Real b, Real *r)
{
    switch (op) {
        case '+': *r = a + b; return 0;
        case '-': *r = a - b; return 0;
        case '*': *r = a * b; return 0;
        case '/': if (0.0 == b) return 1; *r = a / b; return 0;
        case '^': *r = pow(a, b); return 0;
        default: return 1;
    }
}

int artra_expr_eval(const ArtraExpr *e, Real t, Real x, Real y, Real z, Real *ans)
{
    if (e->broken) return 1;
    Real st[MAXSTACK + 1];
    int top = 0;
    for (int q = 0; q < e->n; ++q) {
        const Op *op = e->op + q;
        switch (op->kind) {
            case OP_NUM:
                st[top++] = op->value;
                break;
            case OP_VAR:
                switch (op->sym) {
                    case 'u': st[top++] = t; break;
                    case 'x': st[top++] = x; break;
                    case 'y': st[top++] = y; break;
                    case 'z': st[top++] = z; break;
                    case 'p': st[top++] = 3.14159265358979323846; break;
                    default: st[top++] = *ans; break;
                }
                break;
            case OP_UNARY:
                if (Unary(op->sym, st[top - 1], &st[top - 1])) return 1;
                break;
            case OP_BINARY:
                if (Binary(st[top - 2], op->sym, st[top - 1], &st[top - 2])) return 1;
                --top;
                break;
            default:
                return 1;
        }
    }
    *ans = st[0];
    return 0;
}

int artra_calc(const char *text, Real t, Real x, Real y, Real z, Real *ans)
{
    ArtraExpr *e = NULL;
    if (artra_expr_compile(text, &e)) return 1;
    const int rc = artra_expr_eval(e, t, x, y, z, ans);
    artra_expr_free(e);
    return rc;
}
