// bsg_layout.cpp -- the reference's mask layout, rebuilt bit-exactly.
//
// Follows bsglmm/main.cu:189-394 (6-neighbour dilation mskp, TOTVOXp, the
// mask rank hostIdx and dilated rank hostIdxSC in raster order, the two
// red/black voxel lists with in-mask neighbour counts from cnt_nbrs,
// main.cu:471-502). Adds the colour-sorted order and neighbour table that the
// device kernels use.
#include "bsg_internal.h"

namespace bsg {

int build_layout(const unsigned char *mask, int nrow, int ncol, int ndepth, Layout &L)
{
  if (!mask || nrow < 1 || ncol < 1 || ndepth < 1) {
    set_error("layout: empty grid");
    return -1;
  }
  const int sx = nrow + 2, sxy = (nrow + 2) * (ncol + 2);
  const long long nsize_ll = (long long)sxy * (ndepth + 2);
  if (nsize_ll > (1LL << 30)) {
    set_error("layout: grid too large");
    return -1;
  }
  const int nsize = (int)nsize_ll;
  L.nrow = nrow;
  L.ncol = ncol;
  L.ndepth = ndepth;
  L.nsize = nsize;

  // the padding border is never in the mask (read_data.cpp:231-240 only fills
  // the interior); a caller-supplied grid must respect that
  for (int k = 0; k < ndepth + 2; k++)
    for (int j = 0; j < ncol + 2; j++)
      for (int i = 0; i < nrow + 2; i++) {
        bool border = i == 0 || j == 0 || k == 0 || i == nrow + 1 || j == ncol + 1 || k == ndepth + 1;
        if (border && mask[i + sx * j + sxy * k]) {
          set_error("layout: mask set on the padding border");
          return -1;
        }
      }

  std::vector<unsigned char> mskp((size_t)nsize, 0);
  const int off[6] = {-1, 1, -sx, sx, -sxy, sxy};
  for (int k = 1; k <= ndepth; k++)
    for (int j = 1; j <= ncol; j++)
      for (int i = 1; i <= nrow; i++) {
        int id = i + sx * j + sxy * k;
        if (mask[id]) {
          mskp[id] = 1;
          for (int d = 0; d < 6; d++) mskp[id + off[d]] = 1;
        }
      }

  L.host_idx.assign((size_t)nsize, -1);
  L.host_idx_sc.assign((size_t)nsize, -1);
  int cnt = 0, cntp = 0;
  for (int id = 0; id < nsize; id++) {
    if (mask[id]) L.host_idx[id] = cnt++;
    if (mskp[id]) L.host_idx_sc[id] = cntp++;
  }
  L.totvox = cnt;
  L.totvoxp = cntp;
  if (cnt == 0) {
    set_error("layout: mask is empty");
    return -1;
  }

  for (int c = 0; c < 2; c++) {
    L.vox[c].clear();
    L.nbrs[c].clear();
  }
  for (int k = 1; k <= ndepth; k++)
    for (int j = 1; j <= ncol; j++)
      for (int i = 1; i <= nrow; i++) {
        int id = i + sx * j + sxy * k;
        if (!mask[id]) continue;
        int c = ((i + j + k) & 1) ? 0 : 1;  // main.cu:221-247
        unsigned char m = 0;
        for (int d = 0; d < 6; d++) m += mask[id + off[d]] != 0;
        L.vox[c].push_back(id);
        L.nbrs[c].push_back(m);
      }
  L.n[0] = (int)L.vox[0].size();
  L.n[1] = (int)L.vox[1].size();

  const int V = L.totvox;
  std::vector<int> sorted((size_t)nsize, -1);
  L.rank_of_sorted.resize(V);
  L.sc_of_sorted.resize(V);
  for (int c = 0, s = 0; c < 2; c++)
    for (int q = 0; q < L.n[c]; q++, s++) {
      int id = L.vox[c][q];
      sorted[id] = s;
      L.rank_of_sorted[s] = L.host_idx[id];
      L.sc_of_sorted[s] = L.host_idx_sc[id];
    }
  L.nbr6.resize((size_t)V * 6);
  for (int c = 0, s = 0; c < 2; c++)
    for (int q = 0; q < L.n[c]; q++, s++) {
      int id = L.vox[c][q];
      for (int d = 0; d < 6; d++) L.nbr6[(size_t)s * 6 + d] = mask[id + off[d]] ? sorted[id + off[d]] : -1;
    }
  return 0;
}

}  // namespace bsg
