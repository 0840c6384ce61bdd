// Host analysis of a block pattern for ILU(0): elimination order (natural or
// greedy multicolour), forward/backward levels of the triangular sweeps, and
// the update lists of the factorisation.  See ilu_plan.h.
#include <algorithm>
#include <numeric>
#include <stdexcept>
#include <string>

#include "ilu_plan.h"

namespace hfx {

namespace {

// Greedy colouring of the symmetrised block graph, rows in natural order.
int greedy_colour(int nb, const int* rowptr, const int* col, std::vector<int>& colour) {
  // transpose pattern (column -> rows) for the symmetric neighbourhood
  std::vector<int> tptr(nb + 1, 0), tcol(rowptr[nb]);
  for (int p = 0; p < rowptr[nb]; ++p) ++tptr[col[p] + 1];
  for (int i = 0; i < nb; ++i) tptr[i + 1] += tptr[i];
  std::vector<int> fill(tptr.begin(), tptr.end() - 1);
  for (int i = 0; i < nb; ++i)
    for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) tcol[fill[col[p]]++] = i;
  colour.assign(nb, -1);
  std::vector<int> stamp;  // stamp[c] == i + 1: colour c used by a neighbour of i
  int ncol = 0;
  for (int i = 0; i < nb; ++i) {
    auto mark = [&](int j) {
      if (j == i || colour[j] < 0) return;
      if ((int)stamp.size() <= colour[j]) stamp.resize(colour[j] + 1, 0);
      stamp[colour[j]] = i + 1;
    };
    for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) mark(col[p]);
    for (int p = tptr[i]; p < tptr[i + 1]; ++p) mark(tcol[p]);
    int c = 0;
    while (c < (int)stamp.size() && stamp[c] == i + 1) ++c;
    colour[i] = c;
    ncol = std::max(ncol, c + 1);
  }
  return ncol;
}

// Group rows by level (stable in elimination order within a level).
void group(const std::vector<int>& lev, const std::vector<int>& order, int nlev, std::vector<int>& rows,
           std::vector<int>& off) {
  off.assign(nlev + 1, 0);
  for (int i : order) ++off[lev[i] + 1];
  for (int l = 0; l < nlev; ++l) off[l + 1] += off[l];
  std::vector<int> fill(off.begin(), off.end() - 1);
  rows.assign(order.size(), 0);
  for (int i : order) rows[fill[lev[i]]++] = i;
}

}  // namespace

IluPlan ilu_analyse(int nb, const int* rowptr, const int* col, int ordering) {
  if (nb <= 0) throw std::runtime_error("ILU(0): empty pattern");
  if (ordering != ILU_MULTICOLOUR && ordering != ILU_NATURAL) throw std::runtime_error("ILU(0): unknown ordering");
  IluPlan P;
  P.nb = nb;
  const int nnzb = rowptr[nb];
  for (int i = 0; i < nb; ++i) {
    if (rowptr[i + 1] < rowptr[i]) throw std::runtime_error("ILU(0): row pointers must not decrease");
    P.maxlen = std::max(P.maxlen, rowptr[i + 1] - rowptr[i]);
  }
  for (int p = 0; p < nnzb; ++p)
    if (col[p] < 0 || col[p] >= nb) throw std::runtime_error("ILU(0): column index out of range");

  // elimination order
  std::vector<int> order(nb);
  if (ordering == ILU_NATURAL) {
    P.colour.assign(nb, 0);
    P.ncolours = 1;
    std::iota(order.begin(), order.end(), 0);
  } else {
    P.ncolours = greedy_colour(nb, rowptr, col, P.colour);
    std::vector<int> cnt(P.ncolours + 1, 0);
    for (int i = 0; i < nb; ++i) ++cnt[P.colour[i] + 1];
    for (int c = 0; c < P.ncolours; ++c) cnt[c + 1] += cnt[c];
    for (int i = 0; i < nb; ++i) order[cnt[P.colour[i]]++] = i;
  }
  P.rank.assign(nb, 0);
  for (int r = 0; r < nb; ++r) P.rank[order[r]] = r;
  const std::vector<int>& rank = P.rank;

  // block kinds, diagonal slots, earlier / later lists
  P.kind.assign(nnzb, 0);
  P.dslot.assign(nb, -1);
  P.lptr.assign(nb + 1, 0);
  P.uptr.assign(nb + 1, 0);
  for (int i = 0; i < nb; ++i) {
    std::vector<std::pair<int, int>> lo;  // (rank of column, position)
    for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
      const int j = col[p];
      if (j == i) {
        if (P.dslot[i] >= 0) throw std::runtime_error("ILU(0): duplicate diagonal block in row " + std::to_string(i));
        P.dslot[i] = p - rowptr[i];
        P.kind[p] = 0;
      } else if (rank[j] < rank[i]) {
        P.kind[p] = -1;
        lo.emplace_back(rank[j], p);
      } else {
        P.kind[p] = 1;
        P.upos.push_back(p);
        P.ucol.push_back(j);
      }
    }
    if (P.dslot[i] < 0) throw std::runtime_error("ILU(0): block row " + std::to_string(i) + " has no diagonal block");
    std::sort(lo.begin(), lo.end());
    for (auto& e : lo) {
      P.lpos.push_back(e.second);
      P.lcol.push_back(col[e.second]);
    }
    P.lptr[i + 1] = (int)P.lpos.size();
    P.uptr[i + 1] = (int)P.upos.size();
  }

  // update pairs of each earlier block (i, k): (slot of j in row i, position of (k, j)), j later than k
  std::vector<int> slot_of(nb, -1);
  P.pptr.assign(P.lpos.size() + 1, 0);
  for (int i = 0; i < nb; ++i) {
    for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
      if (slot_of[col[p]] >= 0) throw std::runtime_error("ILU(0): duplicate block in row " + std::to_string(i));
      slot_of[col[p]] = p - rowptr[i];
    }
    for (int l = P.lptr[i]; l < P.lptr[i + 1]; ++l) {
      const int k = P.lcol[l];
      for (int u = P.uptr[k]; u < P.uptr[k + 1]; ++u) {
        const int j = P.ucol[u];
        if (slot_of[j] >= 0) {
          P.pslot.push_back(slot_of[j]);
          P.ppos.push_back(P.upos[u]);
        }
      }
      P.pptr[l + 1] = (int)P.pslot.size();
    }
    for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) slot_of[col[p]] = -1;
  }

  // levels
  P.levf.assign(nb, 0);
  P.levb.assign(nb, 0);
  int nf = 0, nbk = 0;
  for (int r = 0; r < nb; ++r) {
    const int i = order[r];
    int lv = 0;
    for (int l = P.lptr[i]; l < P.lptr[i + 1]; ++l) lv = std::max(lv, P.levf[P.lcol[l]] + 1);
    P.levf[i] = lv;
    nf = std::max(nf, lv + 1);
  }
  for (int r = nb - 1; r >= 0; --r) {
    const int i = order[r];
    int lv = 0;
    for (int u = P.uptr[i]; u < P.uptr[i + 1]; ++u) lv = std::max(lv, P.levb[P.ucol[u]] + 1);
    P.levb[i] = lv;
    nbk = std::max(nbk, lv + 1);
  }
  P.nlev_f = nf;
  P.nlev_b = nbk;
  group(P.levf, order, nf, P.fwd_rows, P.fwd_off);
  group(P.levb, order, nbk, P.bwd_rows, P.bwd_off);
  return P;
}

}  // namespace hfx
