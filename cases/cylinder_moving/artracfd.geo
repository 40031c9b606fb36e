count begin
0
1
count end
polyhedron geometry begin
cylinder.stl
polyhedron geometry end
polyhedron state begin
0, 0, 0, 0.1, 0.8, 0.5, 0, 0, 0, 3, 2.7, -1, 0, 0, 0, 0
0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0
polyhedron state end
polyhedron transform begin
1, 1, 1, 0, 0, 0, 0, 0, 0
polyhedron transform end
