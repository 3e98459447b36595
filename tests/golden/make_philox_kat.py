"""Known-answer vectors for Philox4x32-10 from ROCm's open-source rocRAND.

Compiles a host-only program against /opt/rocm/include/rocrand/
rocrand_philox4x32_10.h (rocRAND's engine, whose ten_rounds() is the
Salmon et al. SC'11 round function) and records ten_rounds(counter, key) for a
set of counters and keys, including the counter layout the IS-SPA engine
uses (index, stream<<24 | draw, step, system).  Output: philox_kat.json.

    python tests/golden/make_philox_kat.py
"""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

SRC = r"""
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>
struct probe : rocrand_device::philox4x32_10_engine {
    uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main() {
    probe p;
    unsigned c0, c1, c2, c3, k0, k1;
    while (scanf("%x %x %x %x %x %x", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        uint4 c = {c0, c1, c2, c3};
        uint2 k = {k0, k1};
        uint4 o = p.run(c, k);
        printf("%08x %08x %08x %08x\n", o.x, o.y, o.z, o.w);
    }
    return 0;
}
"""


def cases():
    out = [
        ([0, 0, 0, 0], [0, 0]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]),
    ]
    seed = 12345
    for idx in (0, 1, 219, 399959):
        for stream, draw in ((0, 0), (0, 3), (1, 0), (1, 1)):
            for step in (0, 1, 999):
                out.append(([idx, (stream << 24) | draw, step, 0], [seed & 0xFFFFFFFF, seed >> 32]))
    out.append(([7, 5, 123456789, 63], [0xDEADBEEF, 0x01234567]))
    return out


def main():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "k.cpp")
        exe = os.path.join(d, "k")
        open(src, "w").write(SRC)
        subprocess.run(["hipcc", "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-O1", "-o", exe, src,
                        "-I/opt/rocm/include"], check=True)
        cs = cases()
        inp = "\n".join(" ".join("%x" % v for v in c + k) for c, k in cs) + "\n"
        res = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout
    rows = []
    for (c, k), line in zip(cs, res.split("\n")):
        rows.append({"ctr": c, "key": k, "out": [int(x, 16) for x in line.split()]})
    with open(os.path.join(HERE, "philox_kat.json"), "w") as fh:
        json.dump({"source": "rocrand_philox4x32_10.h ten_rounds (ROCm 7.2)", "vectors": rows}, fh,
                  indent=1)
    print(f"wrote {len(rows)} vectors")


if __name__ == "__main__":
    main()
