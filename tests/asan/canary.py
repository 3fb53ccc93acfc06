"""Sanitizer canary: writes 16 camera rays into 4-ray arrays through the
oracle.  Run only under tests/asan/run_asan.sh (ASan preloaded, the sanitizer
build of the oracle selected); there the out-of-bounds write must abort the
process with a heap-buffer-overflow report, which shows the build is
instrumented."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
import oracle  # noqa: E402

lib = oracle._load()
f = lambda a: np.ascontiguousarray(a, np.float32)
pos, quat, th, pp = f([0, 0, 0]), f([0, 0, 0, 1]), f([0.5, 0.5]), f([0, 0])
o = np.zeros((4, 3), np.float32)  # room for 4 rays; the call writes a 4x4 image
d = np.zeros((4, 3), np.float32)
lib.bvo_camera_rays_f32(pos.ctypes.data, quat.ctypes.data, th.ctypes.data, pp.ctypes.data, 4, 4, o.ctypes.data,
                        d.ctypes.data)
print("no report: the oracle is not instrumented")
