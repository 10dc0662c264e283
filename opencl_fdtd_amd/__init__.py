"""MI355X-native FDTD time-stepping (drop-in for opencl_fdtd's Simulation).

Exports mirror opencl_fdtd/__init__.py:1-5.
"""
from .simulation import Simulation, type_to_C, Event, DeviceField

version = '0.4'
