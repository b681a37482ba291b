"""MI355X-native implementation of the Channel_Flow dual-time hot path.

The compute lives in libchannel_flow.so (HIP kernels for gfx950 behind the C
ABI of include/channel_flow.h); this package is the thin host mirror of the
reference's call surface (solver.Solver) plus run configuration and the
seeded initial condition (config).
"""
from .config import ChannelConfig, FlowParams  # noqa: F401

__all__ = ["ChannelConfig", "FlowParams"]
