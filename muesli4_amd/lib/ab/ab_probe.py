import ctypes as C, sys, time, os, json
import numpy as np
sys.path.insert(0, os.getcwd())
from muesli4_amd import hip
from muesli4_amd.hip import DeviceArray, call
hip.init(0)
ms = C.c_float(); inb = C.c_int()
out = {"lib": os.environ.get("MSL_HIP_LIB")}
n = 8192 * 8192
a = DeviceArray(n, np.float32); b = DeviceArray(n, np.float32)
for r in (1, 5, 10):
    call("msl_fill_uniform_f32", a.ptr, n, 7, 255.0 / 16777216.0)
    call("msl_stencil2d_mean_iterate_f32", a.ptr, b.ptr, 8192, 8192, r, 0.0, 20, C.byref(inb))
    ts = []
    for _ in range(7):
        call("msl_timer_start"); call("msl_stencil2d_mean_iterate_f32", a.ptr, b.ptr, 8192, 8192, r, 0.0, 20, C.byref(inb)); call("msl_timer_stop", C.byref(ms)); ts.append(ms.value)
    out[f"r{r}_GBs"] = round(8 * n * 20 / (np.median(ts) * 1e-3) / 1e9, 1)
a.free(); b.free()
m = 512 ** 3
a = DeviceArray(m, np.float32); b = DeviceArray(m, np.float32)
res = (C.c_float * 20)()
call("msl_fill_uniform_f32", a.ptr, m, 7, 1.0 / 16777216.0)
call("msl_heat7_iterate_f32", a.ptr, b.ptr, 512, 512, 512, 0.0, 200, 10, res, C.byref(inb))
ts = []
for _ in range(5):
    call("msl_timer_start"); call("msl_heat7_iterate_f32", a.ptr, b.ptr, 512, 512, 512, 0.0, 200, 10, res, C.byref(inb)); call("msl_timer_stop", C.byref(ms)); ts.append(ms.value)
out["heat200_GBs"] = round(8 * m * 200 / (np.median(ts) * 1e-3) / 1e9, 1)
ts = []
for i in range(20):
    src, dst = (a, b) if i % 2 == 0 else (b, a)
    call("msl_timer_start"); call("msl_heat7_f32", dst.ptr, src.ptr, 512, 512, 512, 0.0); call("msl_timer_stop", C.byref(ms)); ts.append(ms.value)
out["heat1_GBs"] = round(8 * m / (np.median(ts) * 1e-3) / 1e9, 1)
print(json.dumps(out), flush=True)
