"""CloverLeaf 3D input decks.

`bm_deck` is apps/c/CloverLeaf_3D/clover.in (the reference's 3D benchmark: a 96^3 mesh on
[0,10]^3, ambient density 0.2 / energy 1.0, a cuboid x in [0,5], y in [0,2], z in [0,4] at
density 1.0 / energy 2.5, dt 0.1, 87 steps, known answer ke 20.0546870878964 from
field_summary.cpp:95) with the mesh size and step count as parameters.
"""

GOLD_KE_BM = 20.0546870878964   # apps/c/CloverLeaf_3D/field_summary.cpp:95 (test_problem 2)


def bm_deck(n=96, steps=87, ny=None, nz=None):
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    return f"""*clover

 state 1 density=0.2 energy=1.0
 state 2 density=1.0 energy=2.5 geometry=cuboid xmin=0.0 xmax=5.0 ymin=0.0 ymax=2.0 zmin=0.0 zmax=4.0

 x_cells={n}
 y_cells={ny}
 z_cells={nz}

 xmin=0.0
 ymin=0.0
 zmin=0.0
 xmax=10.0
 ymax=10.0
 zmax=10.0

 initial_timestep=0.1
 max_timestep=0.1
 end_step={steps}
 end_time=1000.1
 test_problem 2
 checkpoint_frequency=10
 profiler_on=1
*endclover
"""
