"""MI355X (gfx950) D2Q9 lattice-Boltzmann hot path of simuverse's fluid simulator.

``simuverse_amd.lbm`` binds the C ABI of ``include/lbm.h``;
``simuverse_amd.workloads`` builds the seeded lattices the tests and the
benchmark run.
"""
__all__ = ["lbm", "workloads"]
