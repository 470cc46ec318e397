"""Write tests/golden/philox_vectors.json: Philox4x32-10 blocks computed by
rocRAND's implementation (rocrand_philox4x32_10.h, host-callable ten_rounds),
an independent pin of the oracle's and the HIP path's Philox.  Run here once;
the JSON is committed."""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = r'''
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>
struct Probe : rocrand_device::philox4x32_10_engine {
    uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main() {
    const uint32_t cases[][6] = {
        {0u, 0u, 0u, 0u, 0u, 0u},
        {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
        {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
        {0u, 12345u, 8u, 0u, 0x5EED0002u, 0u},
        {7u, 76799u, 10u, 3u, 0x5EED0002u, 0u},
        {1u, 2u, 3u, 4u, 5u, 6u},
    };
    Probe p;
    printf("[");
    for (int i = 0; i < 6; ++i) {
        uint4 c = {cases[i][0], cases[i][1], cases[i][2], cases[i][3]};
        uint2 k = {cases[i][4], cases[i][5]};
        uint4 o = p.run(c, k);
        printf("%s{\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}", i ? ", " : "",
               c.x, c.y, c.z, c.w, k.x, k.y, o.x, o.y, o.z, o.w);
    }
    printf("]\n");
    return 0;
}
'''


def main():
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "p.cpp")
        exe = os.path.join(td, "p")
        open(src, "w").write(SRC)
        subprocess.check_call(["/opt/rocm/bin/hipcc", "-x", "c++", "-O1", "-D__HIP_PLATFORM_AMD__",
                               "-I/opt/rocm/include", src, "-o", exe])
        out = subprocess.check_output([exe]).decode()
    vec = json.loads(out)
    json.dump({"source": "rocRAND rocrand_philox4x32_10.h ten_rounds (ROCm 7.2)", "vectors": vec},
              open(os.path.join(HERE, "philox_vectors.json"), "w"), indent=1)
    print(json.dumps(vec[:3]))


if __name__ == "__main__":
    main()
