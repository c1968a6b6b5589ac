// Host-side CSR pattern construction for FEM assembly.
//
// Reference behaviour: scipy's COO->CSR conversion of the per-element triplets
// (qdsim_cython/solvers/fem_solver.pyx:142-145, fixed_open_system_solver.pyx:259-267)
// and Eigen's setFromTriplets (backend/src/fem.cpp:148-150).  Both give the union of
// the element DOF cliques with sorted column indices and explicit zeros kept; the
// pattern built here is that canonical pattern.
//
// Besides the pattern this file produces the element->nonzero scatter map and its
// inverse, a segmented list of the packed element-matrix entries that land on each
// nonzero, in ascending (element, i, j) order.  The device gather kernel sums each
// segment in that fixed order, so assembly needs no atomics and is bitwise
// reproducible.
#include "qds_common.h"
#include "../../include/qdsim_hip.h"

#include <algorithm>
#include <thread>
#include <vector>

namespace {

struct DofElements {
  std::vector<int64_t> ptr;   // n_dof + 1
  std::vector<int32_t> elem;  // elements touching each dof (sorted)
};

DofElements dof_to_elements(const int32_t* conn, int64_t ne, int npe, int64_t nd) {
  DofElements d;
  d.ptr.assign(nd + 1, 0);
  for (int64_t e = 0; e < ne; ++e)
    for (int a = 0; a < npe; ++a) d.ptr[conn[e * npe + a] + 1]++;
  for (int64_t i = 0; i < nd; ++i) d.ptr[i + 1] += d.ptr[i];
  d.elem.resize(d.ptr[nd]);
  std::vector<int64_t> fill(d.ptr.begin(), d.ptr.end() - 1);
  for (int64_t e = 0; e < ne; ++e)
    for (int a = 0; a < npe; ++a) d.elem[fill[conn[e * npe + a]]++] = static_cast<int32_t>(e);
  return d;
}

int hw_threads() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 1;
  if (n > 16) n = 16;
  return static_cast<int>(n);
}

template <class F>
void parallel_rows(int64_t n, F f) {
  int nt = hw_threads();
  if (n < 4096 || nt == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    int64_t lo = t * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back(f, lo, hi);
  }
  for (auto& x : th) x.join();
}

// columns of row r: union of dofs of the elements touching r
void row_columns(const DofElements& d, const int32_t* conn, int npe, int64_t r,
                 std::vector<int32_t>& cols) {
  cols.clear();
  for (int64_t k = d.ptr[r]; k < d.ptr[r + 1]; ++k) {
    int64_t e = d.elem[k];
    for (int a = 0; a < npe; ++a) cols.push_back(conn[e * npe + a]);
  }
  std::sort(cols.begin(), cols.end());
  cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
}

bool check_conn(const int32_t* conn, int64_t ne, int npe, int64_t nd) {
  for (int64_t i = 0; i < ne * npe; ++i)
    if (conn[i] < 0 || conn[i] >= nd) return false;
  return true;
}

}  // namespace

// Number of nonzeros of the assembled pattern.
QDS_API int qds_pattern_count(const int32_t* conn, int64_t n_elem, int32_t npe, int64_t n_dof,
                              int64_t* nnz_out) {
  QDS_CHECK_ARG(conn && nnz_out && n_elem >= 0 && npe > 0 && n_dof > 0,
                "qds_pattern_count: bad arguments");
  QDS_CHECK_ARG(check_conn(conn, n_elem, npe, n_dof), "qds_pattern_count: dof id out of range");
  DofElements d = dof_to_elements(conn, n_elem, npe, n_dof);
  std::vector<int64_t> cnt(n_dof, 0);
  parallel_rows(n_dof, [&](int64_t lo, int64_t hi) {
    std::vector<int32_t> cols;
    for (int64_t r = lo; r < hi; ++r) {
      row_columns(d, conn, npe, r, cols);
      cnt[r] = static_cast<int64_t>(cols.size());
    }
  });
  int64_t nnz = 0;
  for (auto c : cnt) nnz += c;
  QDS_CHECK_ARG(nnz < (int64_t(1) << 31), "qds_pattern_count: nnz %lld exceeds int32 indexing",
                (long long)nnz);
  *nnz_out = nnz;
  return QDS_OK;
}

// Fill the pattern, the scatter map and its inverse.
//   indptr[n_dof+1], indices[nnz]          canonical CSR pattern
//   elem_map[n_elem*npe*npe] (optional)    nonzero index of element entry (e, i, j)
//   seg_ptr[nnz+1], seg_src[n_elem*npe*npe] packed-symmetric element-entry index
//       e*nsym + s(min(i,j), max(i,j)), s(i,j) = i*npe - i*(i-1)/2 + (j-i),
//       grouped by nonzero in ascending (e, i, j) order.
QDS_API int qds_pattern_fill(const int32_t* conn, int64_t n_elem, int32_t npe, int64_t n_dof,
                             int64_t nnz, int32_t* indptr, int32_t* indices, int32_t* seg_ptr,
                             int32_t* seg_src, int32_t* elem_map) {
  QDS_CHECK_ARG(conn && indptr && indices && seg_ptr && seg_src, "qds_pattern_fill: null buffer");
  QDS_CHECK_ARG(check_conn(conn, n_elem, npe, n_dof), "qds_pattern_fill: dof id out of range");
  const int64_t nent = n_elem * npe * npe;
  const int nsym = npe * (npe + 1) / 2;
  QDS_CHECK_ARG(nent < (int64_t(1) << 31) && n_elem * nsym < (int64_t(1) << 31),
                "qds_pattern_fill: too many element entries for int32 maps");
  DofElements d = dof_to_elements(conn, n_elem, npe, n_dof);

  // pass 1: row lengths
  std::vector<int64_t> len(n_dof);
  parallel_rows(n_dof, [&](int64_t lo, int64_t hi) {
    std::vector<int32_t> cols;
    for (int64_t r = lo; r < hi; ++r) {
      row_columns(d, conn, npe, r, cols);
      len[r] = static_cast<int64_t>(cols.size());
    }
  });
  int64_t acc = 0;
  indptr[0] = 0;
  for (int64_t r = 0; r < n_dof; ++r) {
    acc += len[r];
    indptr[r + 1] = static_cast<int32_t>(acc);
  }
  QDS_CHECK_ARG(acc == nnz, "qds_pattern_fill: nnz mismatch (%lld vs %lld)", (long long)acc,
                (long long)nnz);
  // pass 2: columns
  parallel_rows(n_dof, [&](int64_t lo, int64_t hi) {
    std::vector<int32_t> cols;
    for (int64_t r = lo; r < hi; ++r) {
      row_columns(d, conn, npe, r, cols);
      std::copy(cols.begin(), cols.end(), indices + indptr[r]);
    }
  });
  // scatter map: nonzero position of every element entry
  std::vector<int32_t> pos(nent);
  parallel_rows(n_elem, [&](int64_t lo, int64_t hi) {
    for (int64_t e = lo; e < hi; ++e)
      for (int i = 0; i < npe; ++i) {
        int32_t r = conn[e * npe + i];
        const int32_t* b = indices + indptr[r];
        const int32_t* en = indices + indptr[r + 1];
        for (int j = 0; j < npe; ++j) {
          const int32_t* it = std::lower_bound(b, en, conn[e * npe + j]);
          pos[(e * npe + i) * npe + j] = static_cast<int32_t>(it - indices);
        }
      }
  });
  if (elem_map) std::copy(pos.begin(), pos.end(), elem_map);
  // inverse map by counting sort (stable in (e, i, j) order)
  std::vector<int32_t> cnt(nnz + 1, 0);
  for (int64_t k = 0; k < nent; ++k) cnt[pos[k] + 1]++;
  seg_ptr[0] = 0;
  for (int64_t t = 0; t < nnz; ++t) seg_ptr[t + 1] = seg_ptr[t] + cnt[t + 1];
  std::vector<int32_t> fill(seg_ptr, seg_ptr + nnz);
  for (int64_t e = 0; e < n_elem; ++e)
    for (int i = 0; i < npe; ++i)
      for (int j = 0; j < npe; ++j) {
        int a = std::min(i, j), b = std::max(i, j);
        int s = a * npe - a * (a - 1) / 2 + (b - a);
        seg_src[fill[pos[(e * npe + i) * npe + j]]++] = static_cast<int32_t>(e * nsym + s);
      }
  return QDS_OK;
}

// Pattern of the submatrix on a renumbered subset of DOFs.
// new_of_old[old] = new id in [0, n_new) or -1 (dropped).  The reduced matrix is
// A_r[new_i, new_j] = A[old_i, old_j], rows and columns in the new numbering with
// sorted columns.  src[k] gives the nonzero of the full pattern for reduced nonzero k.
QDS_API int qds_pattern_restrict_count(const int32_t* indptr, const int32_t* indices, int64_t n_dof,
                                       const int32_t* new_of_old, int64_t n_new, int64_t* nnz_out) {
  QDS_CHECK_ARG(indptr && indices && new_of_old && nnz_out, "qds_pattern_restrict_count: null");
  int64_t nnz = 0;
  for (int64_t r = 0; r < n_dof; ++r) {
    if (new_of_old[r] < 0) continue;
    QDS_CHECK_ARG(new_of_old[r] < n_new, "qds_pattern_restrict_count: new id out of range");
    for (int32_t k = indptr[r]; k < indptr[r + 1]; ++k)
      if (new_of_old[indices[k]] >= 0) ++nnz;
  }
  *nnz_out = nnz;
  return QDS_OK;
}

QDS_API int qds_pattern_restrict_fill(const int32_t* indptr, const int32_t* indices, int64_t n_dof,
                                      const int32_t* new_of_old, int64_t n_new, int64_t nnz_r,
                                      int32_t* r_indptr, int32_t* r_indices, int32_t* r_src) {
  QDS_CHECK_ARG(indptr && indices && new_of_old && r_indptr && r_indices && r_src,
                "qds_pattern_restrict_fill: null");
  std::vector<int64_t> old_of_new(n_new, -1);
  for (int64_t r = 0; r < n_dof; ++r)
    if (new_of_old[r] >= 0) {
      QDS_CHECK_ARG(new_of_old[r] < n_new && old_of_new[new_of_old[r]] < 0,
                    "qds_pattern_restrict_fill: new_of_old is not injective");
      old_of_new[new_of_old[r]] = r;
    }
  for (int64_t i = 0; i < n_new; ++i)
    QDS_CHECK_ARG(old_of_new[i] >= 0, "qds_pattern_restrict_fill: new id %lld unused", (long long)i);
  std::vector<int64_t> len(n_new);
  for (int64_t i = 0; i < n_new; ++i) {
    int64_t r = old_of_new[i], c = 0;
    for (int32_t k = indptr[r]; k < indptr[r + 1]; ++k)
      if (new_of_old[indices[k]] >= 0) ++c;
    len[i] = c;
  }
  int64_t acc = 0;
  r_indptr[0] = 0;
  for (int64_t i = 0; i < n_new; ++i) {
    acc += len[i];
    r_indptr[i + 1] = static_cast<int32_t>(acc);
  }
  QDS_CHECK_ARG(acc == nnz_r, "qds_pattern_restrict_fill: nnz mismatch");
  parallel_rows(n_new, [&](int64_t lo, int64_t hi) {
    std::vector<std::pair<int32_t, int32_t>> row;
    for (int64_t i = lo; i < hi; ++i) {
      int64_t r = old_of_new[i];
      row.clear();
      for (int32_t k = indptr[r]; k < indptr[r + 1]; ++k) {
        int32_t nc = new_of_old[indices[k]];
        if (nc >= 0) row.emplace_back(nc, k);
      }
      std::sort(row.begin(), row.end());
      int64_t o = r_indptr[i];
      for (auto& p : row) {
        r_indices[o] = p.first;
        r_src[o] = p.second;
        ++o;
      }
    }
  });
  return QDS_OK;
}

// ---------------------------------------------------------------------------
// Lattice (stencil) form of a lattice-ordered operator: per-class offset lists
// and the map from value slots to CSR positions.  Rows are the nx * ny interior
// lattice points in row-major order; row r has full-lattice coordinates
// I = r % nx + 1, J = r / nx + 1 and class (J % p) * p + I % p.
QDS_API int qds_lattice_stencil_build(int32_t order, int64_t nx, int64_t ny, const int32_t* indptr,
                                      const int32_t* indices, qds_lattice_stencil* st,
                                      int64_t* slot_src) {
  QDS_CHECK_ARG(indptr && indices && st, "qds_lattice_stencil_build: null pointer");
  QDS_CHECK_ARG(order >= 1 && order <= 3, "qds_lattice_stencil_build: order %d not in 1..3", order);
  QDS_CHECK_ARG(nx >= 1 && ny >= 1, "qds_lattice_stencil_build: empty lattice");
  const int p = order, w = 2 * p + 1, nc = p * p;
  const int64_t n = nx * ny;
  // union of the offsets of every row of a class, as a (2p+1)^2 bitmap per class
  const size_t nbits = static_cast<size_t>(nc) * w * w;
  const int nt = hw_threads();
  std::vector<uint8_t> used_t(nbits * nt, 0);
  std::vector<int> bad_t(nt, 0);
  {
    std::vector<std::thread> th;
    const int64_t chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
      th.emplace_back([&, t]() {
        uint8_t* used = used_t.data() + nbits * t;
        const int64_t lo = t * chunk, hi = std::min(n, lo + chunk);
        for (int64_t r = lo; r < hi; ++r) {
          const int64_t i = r % nx, j = r / nx;
          const int c = static_cast<int>(((j + 1) % p) * p + (i + 1) % p);
          for (int32_t k = indptr[r]; k < indptr[r + 1]; ++k) {
            const int64_t col = indices[k];
            const int64_t di = col % nx - i, dj = col / nx - j;
            if (col < 0 || col >= n || di < -p || di > p || dj < -p || dj > p) {
              bad_t[t] = 1;
              return;
            }
            used[(static_cast<size_t>(c) * w + (dj + p)) * w + (di + p)] = 1;
          }
        }
      });
    }
    for (auto& x : th) x.join();
  }
  std::vector<uint8_t> used(nbits, 0);
  int bad = 0;
  for (int t = 0; t < nt; ++t) {
    bad |= bad_t[t];
    for (size_t q = 0; q < nbits; ++q) used[q] |= used_t[nbits * t + q];
  }
  QDS_CHECK_ARG(!bad, "qds_lattice_stencil_build: a column lies more than %d lattice steps from its row", p);
  *st = qds_lattice_stencil{};
  st->order = p;
  st->n_class = nc;
  st->nx = nx;
  st->ny = ny;
  st->cells_x = nx / p + 1;
  st->cells_y = ny / p + 1;
  // tiles of cells along x: one LDS tile of the product kernel holds whole tiles
  st->tile = p == 1 ? 20 : p == 2 ? 10 : 5;
  st->tiles_x = (st->cells_x + st->tile - 1) / st->tile;
  // slot index of each bitmap position within its class (ascending column order:
  // dJ major, dI minor, since |dI| <= p < nx)
  std::vector<int> kidx(used.size(), -1);
  int64_t base = 0;
  for (int c = 0; c < nc; ++c) {
    int len = 0;
    for (int q = 0; q < w * w; ++q) {
      const size_t pos = static_cast<size_t>(c) * w * w + q;
      if (!used[pos]) continue;
      QDS_CHECK_ARG(len < QDS_LATTICE_MAX_ENTRY, "qds_lattice_stencil_build: stencil too long");
      kidx[pos] = len;
      st->di[c][len] = static_cast<int8_t>(q % w - p);
      st->dj[c][len] = static_cast<int8_t>(q / w - p);
      ++len;
    }
    const int padded = (len + 3) & ~3;   // whole 4-entry groups; padding entries (0, 0) with value 0
    QDS_CHECK_ARG(padded <= QDS_LATTICE_MAX_ENTRY, "qds_lattice_stencil_build: stencil too long");
    st->len[c] = padded;
    st->base[c] = base;
    base += (int64_t)st->tile * padded;
  }
  st->tile_slots = base;
  st->n_slots = st->cells_y * st->tiles_x * st->tile_slots;
  if (!slot_src) return QDS_OK;
  parallel_rows(st->n_slots, [&](int64_t lo, int64_t hi) {
    for (int64_t t = lo; t < hi; ++t) slot_src[t] = -1;
  });
  parallel_rows(n, [&](int64_t lo, int64_t hi) {
    for (int64_t r = lo; r < hi; ++r) {
      const int64_t i = r % nx, j = r / nx, I = i + 1, J = j + 1;
      const int c = static_cast<int>((J % p) * p + I % p);
      const int64_t X = I / p, Y = J / p;
      const int64_t row0 = (Y * st->tiles_x + X / st->tile) * st->tile_slots + st->base[c] + (X % st->tile) * 4;
      for (int32_t k = indptr[r]; k < indptr[r + 1]; ++k) {
        const int64_t col = indices[k];
        const int64_t di = col % nx - i, dj = col / nx - j;
        const int e = kidx[(static_cast<size_t>(c) * w + (dj + p)) * w + (di + p)];
        slot_src[row0 + (e / 4) * st->tile * 4 + e % 4] = k;
      }
    }
  });
  return QDS_OK;
}
