"""Writes the CPU-emulation build of the LDS advection kernels (test only):
copies the device headers and fh_advect.hip from flowlab_amd/csrc into
`out`, with every unchecked LDS read bounds-checked against its window, every
clamp-free sampling axis compared with the clamped one, the launches made
serial loops and the workgroup reach test evaluated serially."""
import os
import re
import shutil
import sys

CSRC = os.path.join(os.path.dirname(__file__), "..", "..", "flowlab_amd", "csrc")


def make(out: str) -> None:
    os.makedirs(out, exist_ok=True)
    for f in ("fh_device.hpp", "fh_stencil.hpp", "fh_launch.hpp", "fh_batchdesc.hpp"):
        shutil.copy(os.path.join(CSRC, f), out)
    p = os.path.join(out, "fh_device.hpp")
    dev = open(p).read()
    subs = [
        ("__device__ __forceinline__ float at(int i, int j) const { return l[off(i, j)]; }",
         "__device__ __forceinline__ float at(int i, int j) const { emu_inwin(i - x0, j - y0, 0, "
         "w); return l[off(i, j)]; }"),
        ("    lds_cfp r = l + off(i, j);",
         "    emu_inwin(i - x0, j - y0, 1, w);\n    lds_cfp r = l + off(i, j);"),
        ("    const float fl = floorf(off);\n    return {base + (int)fl, off - fl};",
         "    const float fl = floorf(off);\n    Axis r{base + (int)fl, off - fl};\n"
         "    emu_axis_check(r, clamped_axis(base, off, count));\n    return r;"),
    ]
    for a, b in subs:
        if dev.count(a) != 1:
            raise RuntimeError(f"fh_device.hpp changed: {a[:50]!r} not found once")
        dev = dev.replace(a, b)
    open(p, "w").write(dev)
    src = open(os.path.join(CSRC, "fh_advect.hip")).read()
    src = re.sub(r"(\w+(?:<\w+>)?)<<<(.*?), (dim3\([^)]*\)), 0, s>>>\(", r"emu_launch(\1, \2, \3, ",
                 src, flags=re.S)
    subs = [
        # the two-pass tile bodies
        ("within_reach(m, c, tid)", "emu_within_reach(lds, c, kReach, WN, WW)", 4),
        ("speed ? within_reach_speed(m, sp, c, tid, speed, bid) : "
         "emu_within_reach(lds, c, kReach, WN, WW)",
         "emu_within_reach(lds, c, kReach, WN, WW)", 1),
        # the one-pass dye kernel: its windows, and its ring before it is read
        ("speed ? within_reach_speed<NW>(m, sp, c, tid, speed, bid)\n"
         "                           : within_reach<NW>(m, c, tid)",
         "emu_within_reach(lds, c, kReach, T::WN, WWF)", 1),
        ("    const F h0 = window_as<F>(hat, 0,",
         "    emu_win_h(T::RW) = T::RH;\n    const F h0 = window_as<F>(hat, 0,", 1),
        # the one-pass velocity kernel likewise
        ("  if (!within_reach<NW>(m, c, tid)) {",
         "  if (!emu_within_reach(lds, c, kReach, T::WN, WWF)) {", 1),
        ("    const F ha = window_as<F>(hat, 0,",
         "    emu_win_h(T::RW) = T::RH;\n    const F ha = window_as<F>(hat, 0,", 1),
    ]
    for a, b, k in subs:
        if src.count(a) != k:
            raise RuntimeError(f"fh_advect.hip changed: {a[:60]!r} found {src.count(a)} times")
        src = src.replace(a, b)
    open(os.path.join(out, "fh_advect_emu.cpp"), "w").write('#include "emu.hpp"\n' + src)


if __name__ == "__main__":
    make(sys.argv[1])
