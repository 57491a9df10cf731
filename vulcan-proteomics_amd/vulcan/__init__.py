"""MI355X-native DDA search hot path with the vulcan.gpu / vulcan.index / vulcan.rescore call surface.

Hot path (fragment index build, candidate generation, top-k + pair closure, exact b/y rescoring) runs as
hand-written HIP kernels for gfx950 behind the C ABI declared in include/vulcan_hip.h.
"""
