count begin
0
0
count end
