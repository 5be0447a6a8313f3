"""Python host side of the MI355X GARM-LS per-timestep grid path.

`capi` binds the C-ABI of libgarm_gpu.so; `scenes` holds the grid convention
and the scene builders of the benchmark configurations; `driver` restates the
reference's frame/step time-step rule (Simulator.cpp:102-128).
"""
from . import capi, scenes, driver  # noqa: F401
