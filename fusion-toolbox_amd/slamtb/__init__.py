"""slamtb -- MI355X-native ICP Gauss-Newton step of the fusion-toolbox.

Drop-in for the reference's ``slamtb.registration`` ICP API and the ICP half
of its ``_cslamtb`` native module. All compute runs through the HIP library
``libslamtb_hip.so`` on a HIP device.
"""
__version__ = "0.1.0"
