$ 5 x 5 mm plate at z = 0, 8 triangles with normal -z
BEGIN BULK
GRID         100         -1.13-3 -1.07-3     0.0
GRID         101          1.41-3 -1.07-3     0.0
GRID         102          3.95-3 -1.07-3     0.0
GRID         103         -1.13-3  1.37-3     0.0
GRID         104          1.41-3  1.37-3     0.0
GRID         105          3.95-3  1.37-3     0.0
GRID         106         -1.13-3  4.02-3     0.0
GRID         107          1.41-3  4.02-3     0.0
GRID         108          3.95-3  4.02-3     0.0
CTRIA3         1       1     100     104     101
CTRIA3         2       1     100     103     104
CTRIA3         3       1     101     105     102
CTRIA3         4       1     101     104     105
CTRIA3         5       1     103     107     104
CTRIA3         6       1     103     106     107
CTRIA3         7       1     104     108     105
CTRIA3         8       1     104     107     108
ENDDATA
