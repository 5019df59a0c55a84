$ unit square (2 triangles) at z = 5e-4 and a tilted triangle
BEGIN BULK
GRID          10             0.0     0.0   5.0-4
GRID          11             1.0     0.0   5.0-4
GRID          12             1.0     1.0   5.0-4
GRID          13             0.0     1.0 0.00050
GRID          20          -2.5-1   1.0+0  -1.0+0
GRID          21         2.5E-01     1.0 -1.00D0
GRID          22             0.0     2.0     0.0
CTRIA3         1       1      10      11      12
CTRIA3         2       1      10      12      13
CTRIA3         3       1      20      21      22
ENDDATA
