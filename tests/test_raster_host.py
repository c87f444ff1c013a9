"""CPU checks of the blend kernels' host-compilable geometry (csrc/ubs_raster.h):
the sub-block culling and the workgroup -> tile mapping.

The exact ellipse/rectangle test decides which 8x8 pixel sub-blocks a
primitive is evaluated on.  It must be conservative: no pixel whose fp32
sigma the kernels accept (0 <= sigma < 1) may sit in a culled sub-block.
The header is compiled for the host (hipcc --cuda-host-only) and checked on
random conics spanning the condition numbers the projection can produce
(conic eigenvalues in [4e-4, 3.3]: eps2d = 0.3 bounds them from above),
with the per-pixel sigma computed as the kernels do.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#pragma clang fp contract(off)
#include "ubs_raster.h"
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
int main(int argc, char **argv) {
    const int n = atoi(argv[1]);
    std::mt19937_64 rng(atoi(argv[2]));
    std::uniform_real_distribution<double> U(0, 1);
    long long tested = 0, culled = 0, bad = 0;
    const double lo = log(4e-4), hi = log(3.3);
    for (int k = 0; k < n; ++k) {
        const double l1 = exp(lo + U(rng) * (hi - lo)), l2 = exp(lo + U(rng) * (hi - lo));
        const double th = U(rng) * M_PI, cs = cos(th), sn = sin(th);
        const float a = (float)(l1 * cs * cs + l2 * sn * sn), c = (float)(l1 * sn * sn + l2 * cs * cs);
        const float b = (float)((l1 - l2) * cs * sn);
        const float mx = (float)(U(rng) * 2000 - 40), my = (float)(U(rng) * 1100 - 40);
        const float det = a * c - b * b;
        if (!(det > 0)) continue;
        const float hx = sqrtf(c / det) * 1.001f + 1e-3f, hy = sqrtf(a / det) * 1.001f + 1e-3f;
        const int bx0 = (int)floorf((mx - hx) / 8) - 1, bx1 = (int)ceilf((mx + hx) / 8) + 1;
        const int by0 = (int)floorf((my - hy) / 8) - 1, by1 = (int)ceilf((my + hy) / 8) + 1;
        for (int by = by0; by < by1; ++by)
            for (int bx = bx0; bx < bx1; ++bx) {
                const float lo_x = bx * 8 + 0.5f, lo_y = by * 8 + 0.5f;
                ++tested;
                if (ubs::ellipse_meets_rect(mx, my, a, b, c, lo_x, lo_x + 7.f, lo_y, lo_y + 7.f)) continue;
                ++culled;
                for (int l = 0; l < 64; ++l) {
                    const float dx = mx - (lo_x + (float)(l & 7)), dy = my - (lo_y + (float)(l >> 3));
                    const float s = fmaf(2.f * b * dx, dy, fmaf(a * dx, dx, (c * dy) * dy));
                    if (s >= 0.f && s < 1.f) ++bad;
                }
            }
    }
    printf("%lld %lld %lld\n", tested, culled, bad);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc not on PATH")
def test_subblock_culling_is_conservative(tmp_path):
    src = tmp_path / "drv.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "drv"
    subprocess.run(["hipcc", "--cuda-host-only", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "universal-beta-splatting_amd", "csrc"), str(src), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    for seed in (1, 2):
        out = subprocess.run([str(exe), "60000", str(seed)], check=True, capture_output=True, text=True, timeout=300)
        tested, culled, bad = map(int, out.stdout.split())
        assert bad == 0, f"{bad} accepted pixels in culled sub-blocks"
        assert culled > 0.5 * tested  # the test has teeth: most candidate sub-blocks are culled


MAP_DRIVER = r"""
#include "ubs_raster.h"
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
// argv: C tw th [seed]: with a seed, each XCD runs its own tiles in a random
// order (as raster_block_order_kernel writes them heaviest first), and an
// order of nothing but padding must give no tile
int main(int argc, char **argv) {
    const uint32_t C = atoi(argv[1]), tw = atoi(argv[2]), th = atoi(argv[3]);
    const uint32_t total = C * tw * th, grid = ubs::xcd_grid(C, tw, th);
    // per XCD, the tiles raster order deals it, shuffled and padded with ~0u
    const uint32_t slots = ubs::xcd_slots(C, tw, th);
    std::vector<uint32_t> order((size_t)ubs::kNumXcd * slots, ~0u);
    const bool ordered = argc > 4;
    if (ordered) {
        std::vector<std::vector<uint32_t>> mine(ubs::kNumXcd);
        for (uint32_t b = 0; b < grid; ++b) {
            uint32_t t;
            if (!ubs::xcd_tile(b, C, tw, th, t)) continue;
            if (ubs::xcd_of_tile(t, C, tw, th) != b % ubs::kNumXcd) { printf("xcd_of_tile disagrees\n"); return 1; }
            mine[b % ubs::kNumXcd].push_back(t);
        }
        std::mt19937 rng(atoi(argv[4]));
        for (int x = 0; x < ubs::kNumXcd; ++x) {
            if (mine[x].size() > slots) { printf("xcd %d holds %zu > %u tiles\n", x, mine[x].size(), slots); return 1; }
            std::shuffle(mine[x].begin(), mine[x].end(), rng);
            for (size_t j = 0; j < mine[x].size(); ++j) order[(size_t)x * slots + j] = mine[x][j];
        }
        uint32_t t;
        std::vector<uint32_t> pad((size_t)ubs::kNumXcd * slots, ~0u);  // only padding: no tile at all
        for (uint32_t b = 0; b < grid; ++b)
            if (ubs::xcd_tile(b, C, tw, th, t, pad.data())) { printf("accepted a padding entry\n"); return 1; }
    }
    std::vector<int> seen(total, 0), xcd(total, -1);
    for (uint32_t b = 0; b < grid; ++b) {
        uint32_t t;
        if (!ubs::xcd_tile(b, C, tw, th, t, ordered ? order.data() : nullptr)) continue;
        if (t >= total) { printf("out of range %u\n", t); return 1; }
        seen[t] += 1;
        xcd[t] = b % ubs::kNumXcd;
    }
    for (uint32_t t = 0; t < total; ++t)
        if (seen[t] != 1) { printf("tile %u covered %d times\n", t, seen[t]); return 1; }
    for (uint32_t t = 0; t < total; ++t) printf("%d ", xcd[t]);
    printf("\n");
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc not on PATH")
def test_workgroup_tile_map_is_a_balanced_bijection(tmp_path):
    """Every (camera, tile) is blended by exactly one workgroup of the launch
    grid, in raster block order and in any other order of the blocks (an
    order entry past the block count is refused), and a centre-heavy image
    spreads evenly over the 8 XCDs."""
    import numpy as np

    src = tmp_path / "map.cpp"
    src.write_text(MAP_DRIVER)
    exe = tmp_path / "map"
    subprocess.run(["hipcc", "--cuda-host-only", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "universal-beta-splatting_amd", "csrc"), str(src), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    for C, tw, th in ((1, 120, 68), (1, 1, 1), (3, 5, 7), (2, 85, 64), (1, 16, 16), (8, 240, 135)):
        out = subprocess.run([str(exe), str(C), str(tw), str(th)], capture_output=True, text=True, timeout=300)
        assert out.returncode == 0, out.stdout
        for seed in (1, 2):  # blocks dealt in another order (heaviest first on the GPU): still a bijection
            o2 = subprocess.run([str(exe), str(C), str(tw), str(th), str(seed)], capture_output=True, text=True,
                                timeout=300)
            assert o2.returncode == 0, o2.stdout
        if (C, tw, th) == (1, 120, 68):
            xcd = np.array(out.stdout.split(), dtype=np.int64).reshape(th, tw)
            yy, xx = np.mgrid[0:th, 0:tw]
            work = np.exp(-(((xx - tw / 2) / (tw / 5)) ** 2 + ((yy - th / 2) / (th / 5)) ** 2))  # centre-heavy
            per = np.bincount(xcd.ravel(), weights=work.ravel(), minlength=8)
            assert per.max() / per.mean() < 1.1, per
