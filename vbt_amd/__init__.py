"""vbt_amd -- MI355X-native VBTPACK4 random-access decoder.

The decode runs in libvbt_hip.so (hand-written HIP kernels for gfx950) behind
the C-ABI in include/vbt_hip.h; vbt_amd.hip is its Python host side and
vbt_amd.synth the synthetic-asset tooling.
"""
__all__ = ["hip", "synth"]
