count begin
1
0
count end
sphere state begin
0.05, 0, 0, 0.2, -0.6, 0.2, 0.1, 0, 0, 2, 2.7, 0.0038, 1, 0, 0, 0
0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0
sphere state end
