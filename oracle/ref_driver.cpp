// Driver around the reference's own sampler and camera headers, compiled
// directly from their sources under /root/reference (sobol.h, camera.h,
// helper_math.h).  It writes input/output vectors that pin the oracle's
// restatement of those two headers.  TEST INFRASTRUCTURE ONLY: built into
// oracle/_ref/ by oracle/Makefile, never shipped in the product.
//
// usage:
//   ref_driver sobol  <in.bin> <out.bin>   in: u32 quads (index,p0,p1,p2); out: f32 pairs
//   ref_driver camera <in.bin> <out.bin>   in: f32 records (eye3 center3 up3 fovy w h u v)
//                                          out: f32 records (origin3 right3 up3 ro3 rd3)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include "sobol.h"
#include "camera.h"

static std::vector<unsigned char> slurp(const char* path)
{
    std::vector<unsigned char> b;
    FILE* f = fopen(path, "rb");
    if (!f) { perror(path); exit(2); }
    unsigned char buf[65536];
    size_t k;
    while ((k = fread(buf, 1, sizeof(buf), f)) > 0) b.insert(b.end(), buf, buf + k);
    fclose(f);
    return b;
}

int main(int argc, char** argv)
{
    if (argc != 4) { fprintf(stderr, "usage: %s sobol|camera in out\n", argv[0]); return 2; }
    std::vector<unsigned char> in = slurp(argv[2]);
    FILE* out = fopen(argv[3], "wb");
    if (!out) { perror(argv[3]); return 2; }
    if (!strcmp(argv[1], "sobol")) {
        size_t n = in.size() / 16;
        for (size_t i = 0; i < n; i++) {
            uint32_t q[4];
            memcpy(q, &in[16 * i], 16);
            float xy[2];
            sobol::shuffled_scrambled_sobol_2d(&xy[0], &xy[1], q[0], q[1], q[2], q[3]);
            fwrite(xy, 4, 2, out);
        }
    } else if (!strcmp(argv[1], "camera")) {
        size_t n = in.size() / (14 * 4);
        for (size_t i = 0; i < n; i++) {
            float r[14];
            memcpy(r, &in[56 * i], 56);
            RayGenerator g;
            g.lookat({r[0], r[1], r[2]}, {r[3], r[4], r[5]}, {r[6], r[7], r[8]}, r[9], (int)r[10], (int)r[11]);
            float3 ro, rd;
            g.shoot(&ro, &rd, r[12], r[13]);
            float o[15] = {g.m_origin.x, g.m_origin.y, g.m_origin.z, g.m_right.x, g.m_right.y, g.m_right.z,
                           g.m_up.x,     g.m_up.y,     g.m_up.z,     ro.x,        ro.y,        ro.z,
                           rd.x,         rd.y,         rd.z};
            fwrite(o, 4, 15, out);
        }
    } else {
        fprintf(stderr, "unknown mode %s\n", argv[1]);
        return 2;
    }
    fclose(out);
    return 0;
}
