PGEOMETRY V5
NPoints 5 NPrims 1
NPointGroups 0 NPrimGroups 0
NPointAttrib 2 NVertexAttrib 0 NPrimAttrib 1 NAttrib 0
PointAttrib

life 2 float 0 0
id 1 int 0
0 0.1 0.2 1 (-1.2 10	0)
0.1 0.2 0.3 1 (-0.2 10	1)
0.2 0.3 0.4 1 (0.8 10	2)
0.3 0.4 0.5 1 (1.8 10	3)
0.4 0.5 0.6 1 (2.8 10	4)
PrimitiveAttrib
generator 1 index 1 papi
Part 5 0 1 2 3 4 [0]
beginExtra
endExtra
