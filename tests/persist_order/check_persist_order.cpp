// Host check of the persistent op's work-item order (vfs::persist_rank, vfs_engine.h) for
// every tile shape that has the op: a bijection of the staged region, the tile's border
// (columns 0, tw - 2, tw - 1 and rows 0, th - 2, th - 1) first and the rest after it, each
// in raster order, and the border within the first position of every work-item.  Prints
// one line per shape and exits non-zero on any violation.
#include "../../vfs_amd/csrc/vfs_engine.h"

#include <cstdio>
#include <vector>

int main()
{
    using namespace vfs;
    int bad = 0, shapes = 0;
    for (int sh = 0; sh < kShapes; sh++) {
        if (!persist_shape(sh)) continue;
        shapes++;
        const int tw = kShape[sh].tw, th = kShape[sh].th, rw = tw + 3, rh = th + 3, nr = rw * rh;
        const int nb = persist_border(sh);
        const int nt = nr >= 1024 ? 1024 : (nr + 63) / 64 * 64;  // work-items of the launch
        std::vector<int> seen(nr, 0);
        int last_border = -1, last_rest = nb - 1;
        for (int i = 0; i < nr; i++) {
            const int q = persist_rank(sh, i);
            if (q < 0 || q >= nr) {
                std::printf("shape %d: raster index %d has rank %d outside [0, %d)\n", sh, i, q, nr);
                bad++;
                continue;
            }
            seen[q]++;
            const int dx = i % rw - 2, dy = i / rw - 2;
            const bool border = dx >= 0 && dx < tw && dy >= 0 && dy < th &&
                                (dx == 0 || dx == tw - 2 || dx == tw - 1 || dy == 0 || dy == th - 2 || dy == th - 1);
            if (border != (q < nb)) {
                std::printf("shape %d: position (%d, %d) has rank %d (border %d)\n", sh, dx, dy, q, (int)border);
                bad++;
            }
            int &last = border ? last_border : last_rest;
            if (q != last + 1) {
                std::printf("shape %d: position (%d, %d) has rank %d after %d (not raster order)\n", sh, dx, dy, q,
                            last);
                bad++;
            }
            last = q;
        }
        for (int q = 0; q < nr; q++)
            if (seen[q] != 1) {
                std::printf("shape %d: rank %d taken %d times\n", sh, q, seen[q]);
                bad++;
            }
        if (nb > nt) {
            std::printf("shape %d: %d border positions for %d work-items\n", sh, nb, nt);
            bad++;
        }
        std::printf("shape %d (%dx%d): %d staged positions, %d on the border, %d work-items\n", sh, tw, th, nr, nb, nt);
    }
    std::printf("%d shapes, %d violations\n", shapes, bad);
    return bad != 0;
}
