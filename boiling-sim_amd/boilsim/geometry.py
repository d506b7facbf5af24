"""Material ids and the grid contract of the reference (python/boilingsim/
geometry.py:39-42, 254-282).

Cell fields are C-contiguous (nx, ny, nz) with k fastest; faces are
ux (nx+1,ny,nz), uy (nx,ny+1,nz), uz (nx,ny,nz+1).
"""

MAT_FLUID = 0
MAT_POT_WALL = 1
MAT_AIR = 2
MAT_CARROT = 3
