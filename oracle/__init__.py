"""CPU oracle for the FDTD(+PIC) time-step loop.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the `py-fdtd_pic_amd`
package, the C-ABI library, the drop-in modules) may import, link or call
anything in this directory.  Only `tests/`, `__graft_entry__.smoke()` and the
`cpu_baseline` leg of `bench.py` use it, and only as the checker or as the
timed CPU baseline, never as the thing measured or shipped.

Modules
-------
faithful   NumPy FP64 restatement of the reference integrators
           (Solver_Engine.py / BaseFDTD11.py / Environment_Setup.py /
           Material_Def.py / genericStability.py / MasterController.py),
           including the reference's float32 fields and coefficient quirks.
cubic      Restatement of CubicEquationSolver.solve (1728.org branches) and
           of the reference's root selection (BaseFDTD11.py:824-853).
physical   NumPy FP64 model used for the configs the reference cannot run
           (soft Gaussian source, eps_inf, per-cell Lorentz, Kerr, CPML,
           two-pass reflection protocol, FP64 DFT spectra).
pic        Textbook relativistic Boris push, linear gather and CIC deposit
           (no reference implementation exists: Solver_Engine.py:343 TODO).
cpu_ref    ctypes loader for fdtd_cpu.c, the C restatement of the physical
           loop that is timed as the CPU baseline.
"""
