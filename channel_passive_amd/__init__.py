"""MI355X-native RK3-CN time step of the PoongBack channel DNS with a
passive scalar.  The compute path is the HIP library built from
``channel_passive_amd/csrc`` (``libpoongback_hip.so``), reached through the
C ABI declared in ``include/pb_capi.h``; see ``channel_passive_amd.capi``."""

from .config import ChannelConfig, PbConfig, preset  # noqa: F401
