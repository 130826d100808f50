"""MI355X-native RK3/Crank-Nicolson substep of the rough-channel DNS.

The compute path is libcfd_amd.so (hand-written HIP kernels for gfx950 behind
the C ABI of include/cfd_amd.h); this package is its host-side mirror.
"""
from .params import Params
from .channel import Channel, build_ygrid, build_iblist, FIELDS, SCALARS, KCLASSES
from . import ic, workloads

__all__ = ["Params", "Channel", "build_ygrid", "build_iblist", "FIELDS", "SCALARS", "KCLASSES", "ic", "workloads"]
