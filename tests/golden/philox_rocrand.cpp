// Writes Philox4x32-10 outputs computed by rocRAND's own engine
// (rocrand_philox4x32_10.h, host path) for a set of counters and keys, as JSON.
// Built and run by tests/golden/make_philox_rocrand.py; the output is the
// committed fixture tests/golden/philox_rocrand.json.
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>

struct Exposed : rocrand_device::philox4x32_10_engine {
    uint4 run(uint4 c, uint2 k) { return this->ten_rounds(c, k); }
};

int main()
{
    Exposed e;
    unsigned s = 12345u;
    auto nxt = [&]() { s = s * 1664525u + 1013904223u; return s; };
    std::printf("{\"source\": \"rocRAND philox4x32_10_engine::ten_rounds\", \"vectors\": [\n");
    for (int i = 0; i < 64; i++) {
        uint4 c;
        uint2 k;
        if (i == 0) { c = {0, 0, 0, 0}; k = {0, 0}; }
        else if (i == 1) { c = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu}; k = {0xffffffffu, 0xffffffffu}; }
        else { c = {nxt(), nxt(), nxt(), nxt()}; k = {nxt(), nxt()}; }
        uint4 o = e.run(c, k);
        std::printf("  {\"counter\": [%u, %u, %u, %u], \"key\": [%u, %u], \"output\": [%u, %u, %u, %u]}%s\n",
                    c.x, c.y, c.z, c.w, k.x, k.y, o.x, o.y, o.z, o.w, i == 63 ? "" : ",");
    }
    std::printf("]}\n");
    return 0;
}
