"""MI355X-native FDTD(+PIC) step loop: host engine over libfdtdpic.so.

The package directory `py-fdtd_pic_amd/` also holds flat drop-in modules
named like the reference's (Solver_Engine, BaseFDTD11, BaseFDTD,
MasterController, Environment_Setup, Material_Def, genericStability,
TransformHandler); put that directory on sys.path instead of the
reference's.
"""
__version__ = "0.1.0"
