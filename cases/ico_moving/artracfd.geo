count begin
0
1
count end
polyhedron geometry begin
ico.stl
polyhedron geometry end
polyhedron state begin
0, 0, 0, 0.21, 0.5, -0.3, 0.2, 1.5, -2.0, 2.5, 2.0, -1, 0, 0, 0, 0
0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0
polyhedron state end
polyhedron transform begin
1, 1, 1, 0, 0, 0, 0, 0, 0
polyhedron transform end
