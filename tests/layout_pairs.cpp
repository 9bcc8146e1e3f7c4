// Test helper (built by tests/test_cpu_layout.py from the library's own
// layout.cpp): enumerates every non-excluded pair with r^2 <= rc^2 covered by
// the tile-list skeleton of a (batched) PatchLayout, with the pair kernel's
// arithmetic (FP32 patch-centred coordinates plus FP32 shift, exact r^2) and
// the exclusion bit table, and writes the pairs and the effective positions.
// usage: layout_pairs in.bin rc patchDim pairs.out effpos.out
// A negative system count in in.bin means one system in a general cell: the
// 3 box doubles are then 9 (lattice rows a, b, c) and PatchLayout::buildCell
// lays it out.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <string>
#include <set>
#include "layout.h"
using namespace mi355nb;
int main(int argc, char** argv) {
  FILE* f = fopen(argv[1], "rb");
  int nsys = 0;
  if (fread(&nsys, 4, 1, f) != 1) return 2;
  const bool cellMode = nsys < 0;
  if (cellMode) nsys = 1;
  std::vector<int> start(nsys + 1);
  if (fread(start.data(), 4, nsys + 1, f) != (size_t)(nsys + 1)) return 2;
  int n = start[nsys];
  const int nbox = cellMode ? 9 : 3 * nsys;
  std::vector<double> boxes(nbox);
  if (fread(boxes.data(), 8, nbox, f) != (size_t)nbox) return 2;
  std::vector<double> pos(3 * n);
  if (fread(pos.data(), 8, 3 * n, f) != (size_t)(3 * n)) return 2;
  std::vector<long long> ptr(n + 1);
  if (fread(ptr.data(), 8, n + 1, f) != (size_t)(n + 1)) return 2;
  std::vector<int> idx(ptr[n] > 0 ? ptr[n] : 1);
  if (ptr[n] > 0 && fread(idx.data(), 4, ptr[n], f) != (size_t)ptr[n]) return 2;
  fclose(f);
  double rc = atof(argv[2]), pl = atof(argv[3]);
  std::vector<int64_t> eptr(ptr.begin(), ptr.end());
  std::vector<I2> sm; std::vector<uint32_t> bits; int nl = 0; std::string err;
  if (!buildExclusions(n, eptr.data(), idx.data(), sm, bits, nl, err)) { printf("excl err %s\n", err.c_str()); return 1; }
  PatchLayout L;
  const bool ok = cellMode ? L.buildCell(n, pos.data(), boxes.data(), pl, err)
                           : L.buildBatch(nsys, start.data(), pos.data(), boxes.data(), pl, err);
  if (!ok) { printf("layout err %s\n", err.c_str()); return 1; }
  std::vector<double> q(n, 0.0); std::vector<int> t(n, 0);
  L.pack(pos.data(), q.data(), t.data(), sm.data());
  FILE* o = fopen(argv[4], "wb");
  long long npairs = 0;
  double c2 = rc * rc;
  for (const TileList& tl : L.tileListSkeleton) {
    const PatchRecord& A = L.patches[tl.patchInd[0]];
    const PatchRecord& B = L.patches[tl.patchInd[1]];
    for (int ii = 0; ii < 32; ++ii) {
      int si = tl.iatomStart + ii;
      if (si >= A.atomStart + A.numAtoms) continue;
      float xi = L.xyzq[si].x + tl.offx, yi = L.xyzq[si].y + tl.offy, zi = L.xyzq[si].z + tl.offz;
      int jb = tl.self ? tl.iatomStart : B.atomStart;
      for (int sj = jb; sj < B.atomStart + B.numAtoms; ++sj) {
        if (tl.self && sj <= si) continue;
        float dx = L.xyzq[sj].x - xi, dy = L.xyzq[sj].y - yi, dz = L.xyzq[sj].z - zi;
        double r2 = (double)dx * dx + (double)dy * dy + (double)dz * dz;
        if (r2 > c2) continue;
        int a = L.atomIndex[si], b = L.atomIndex[sj];
        if (isExcluded(L.exclStartMaxdiff[si], bits.data(), a, b)) continue;
        int lo = a < b ? a : b, hi = a < b ? b : a;
        fwrite(&lo, 4, 1, o); fwrite(&hi, 4, 1, o);
        ++npairs;
      }
    }
  }
  fclose(o);
  printf("pairs %lld lists %zu slots %d\n", npairs, L.tileListSkeleton.size(), L.atomStorageSize);
  // effective positions out
  FILE* e = fopen(argv[5], "wb"); fwrite(L.effPos.data(), 8, L.effPos.size(), e); fclose(e);
  return 0;
}
