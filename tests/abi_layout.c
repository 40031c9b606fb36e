/* abi_layout.c - prints the layout of this library's drop-in structures
 * (include/artra_types.h) in the format of the reference harness's layout
 * dump; compiled and run by tests/test_host.py. */
#include <stddef.h>
#include <stdio.h>

#include "artra_types.h"
#include "abi_fields.h"

int main(void)
{
    ABI_LIST
    return 0;
}
