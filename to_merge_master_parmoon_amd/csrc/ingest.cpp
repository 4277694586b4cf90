// ingest.cpp - the on-disk form of one multigrid level at the library boundary.
//
// A ParMooN build holds each level in TSquareStructure3D / TStructure3D CSR arrays
// (RowPtr, KCol: include/FE/Structure.h:19-45, include/FE/Matrix.h:91-129), the Entries
// arrays of the nine A, three B and three BT matrices, the GlobalNumbers / BeginIndex
// arrays of the velocity and pressure spaces and, for a refined level, the parent/child
// relation of its cells.  Its SaveData / ReadData (src/General/MainUtilities.C:3360-3440)
// write and read double arrays as raw native-endian bytes, one array after the other.
// A level dump is a directory holding each of those arrays in a raw file of its own
// (doubles as SaveData writes them, int arrays as raw 32-bit ints) and a manifest.txt with
// the sizes ("key value" per line), so a ParMooN build can write one with SaveData-style
// raw writes of the arrays it already has, and pmg_level_read turns it into the
// pmg_level_desc that pmg_level_add takes.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "pmg.h"

namespace {

struct Dump {
  pmg_level_desc d{};
  std::vector<int> a_rowptr, a_kcol, b_rowptr, b_kcol, bt_rowptr, bt_kcol;
  std::vector<double> a[9], b[3], bt[3];
  std::vector<int> u_begin, u_global, p_begin, p_global, parent, child, ptr_colors, cell_reorder;
  std::vector<double> qq_u, qq_p;
};

const char *kA[9] = {"A11", "A12", "A13", "A21", "A22", "A23", "A31", "A32", "A33"};
const char *kB[3] = {"B1", "B2", "B3"};
const char *kBT[3] = {"B1T", "B2T", "B3T"};

template <typename T>
bool read_raw(const std::string &path, std::vector<T> &v, size_t n) {
  v.resize(n);
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  f.read(reinterpret_cast<char *>(v.data()), std::streamsize(sizeof(T) * n));
  if (size_t(f.gcount()) != sizeof(T) * n) return false;
  f.peek();
  return f.eof();   // the file holds exactly n values
}

template <typename T>
bool write_raw(const std::string &path, const T *p, size_t n) {
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) return false;
  if (n) f.write(reinterpret_cast<const char *>(p), std::streamsize(sizeof(T) * n));
  return bool(f);
}

}  // namespace

extern "C" {

int pmg_level_read(const char *dir, pmg_level_desc *out, void **handle) {
  if (!dir || !out || !handle) return PMG_ERR_INVALID_ARGUMENT;
  *handle = nullptr;
  const std::string base(dir);
  std::ifstream mf(base + "/manifest.txt");
  if (!mf) return PMG_ERR_INVALID_ARGUMENT;
  std::map<std::string, long> m;
  std::string line;
  while (std::getline(mf, line)) {
    std::istringstream ls(line);
    std::string k;
    long v;
    if (ls >> k >> v) m[k] = v;
  }
  for (const char *k : {"n_udof", "n_active", "hanging_bound", "n_dirichlet", "n_pdof", "a_nnz", "b_nnz", "bt_nnz",
                        "n_cells", "n_ulocal", "n_plocal"})
    if (!m.count(k) || m[k] < 0) return PMG_ERR_INVALID_ARGUMENT;
  auto D = std::make_unique<Dump>();
  pmg_level_desc &d = D->d;
  d.n_udof = int(m["n_udof"]);
  d.n_active = int(m["n_active"]);
  d.hanging_bound = int(m["hanging_bound"]);
  d.n_dirichlet = int(m["n_dirichlet"]);
  d.n_pdof = int(m["n_pdof"]);
  d.a_nnz = int(m["a_nnz"]);
  d.b_nnz = int(m["b_nnz"]);
  d.bt_nnz = int(m["bt_nnz"]);
  d.n_cells = int(m["n_cells"]);
  d.n_ulocal = int(m["n_ulocal"]);
  d.n_plocal = int(m["n_plocal"]);
  d.n_children = m.count("n_children") ? int(m["n_children"]) : 0;
  d.n_colors = m.count("n_colors") ? int(m["n_colors"]) : 0;
  const size_t nu = size_t(d.n_udof), np = size_t(d.n_pdof), nc = size_t(d.n_cells);
  bool ok = read_raw(base + "/a_rowptr.i32", D->a_rowptr, nu + 1) && read_raw(base + "/a_kcol.i32", D->a_kcol, d.a_nnz) &&
            read_raw(base + "/b_rowptr.i32", D->b_rowptr, np + 1) && read_raw(base + "/b_kcol.i32", D->b_kcol, d.b_nnz) &&
            read_raw(base + "/bt_rowptr.i32", D->bt_rowptr, nu + 1) &&
            read_raw(base + "/bt_kcol.i32", D->bt_kcol, d.bt_nnz);
  for (int k = 0; ok && k < 9; ++k) ok = read_raw(base + "/" + kA[k] + ".f64", D->a[k], d.a_nnz);
  for (int k = 0; ok && k < 3; ++k)
    ok = read_raw(base + "/" + kB[k] + ".f64", D->b[k], d.b_nnz) &&
         read_raw(base + "/" + kBT[k] + ".f64", D->bt[k], d.bt_nnz);
  // BeginIndex holds n_cells + 1 entries (the last one ends the GlobalNumbers array)
  ok = ok && read_raw(base + "/u_begin.i32", D->u_begin, nc + 1) && read_raw(base + "/p_begin.i32", D->p_begin, nc + 1);
  ok = ok && D->u_begin[nc] >= 0 && D->p_begin[nc] >= 0 &&
       read_raw(base + "/u_global.i32", D->u_global, size_t(D->u_begin[nc])) &&
       read_raw(base + "/p_global.i32", D->p_global, size_t(D->p_begin[nc]));
  if (ok && d.n_children > 0) {
    const size_t nul = size_t(d.n_ulocal), npl = size_t(d.n_plocal), nch = size_t(d.n_children);
    ok = read_raw(base + "/parent.i32", D->parent, nc) && read_raw(base + "/child.i32", D->child, nc) &&
         read_raw(base + "/qq_u.f64", D->qq_u, nch * nul * nul) && read_raw(base + "/qq_p.f64", D->qq_p, nch * npl * npl);
  }
  if (ok && d.n_colors > 0)
    ok = read_raw(base + "/ptr_colors.i32", D->ptr_colors, size_t(d.n_colors) + 1) &&
         read_raw(base + "/cell_reorder.i32", D->cell_reorder, nc);
  if (!ok) return PMG_ERR_INVALID_ARGUMENT;
  // the arrays must describe CSR matrices and cell maps of the stated sizes
  auto csr_ok = [](const std::vector<int> &rp, const std::vector<int> &kc, long nnz, long ncols) {
    if (rp.front() != 0 || rp.back() != nnz) return false;
    for (size_t i = 0; i + 1 < rp.size(); ++i)
      if (rp[i + 1] < rp[i]) return false;
    for (int c : kc)
      if (c < 0 || c >= ncols) return false;
    return true;
  };
  if (!csr_ok(D->a_rowptr, D->a_kcol, d.a_nnz, d.n_udof) || !csr_ok(D->b_rowptr, D->b_kcol, d.b_nnz, d.n_udof) ||
      !csr_ok(D->bt_rowptr, D->bt_kcol, d.bt_nnz, d.n_pdof))
    return PMG_ERR_INVALID_ARGUMENT;
  d.a_rowptr = D->a_rowptr.data();
  d.a_kcol = D->a_kcol.data();
  for (int k = 0; k < 9; ++k) d.a[k] = D->a[k].data();
  d.b_rowptr = D->b_rowptr.data();
  d.b_kcol = D->b_kcol.data();
  d.bt_rowptr = D->bt_rowptr.data();
  d.bt_kcol = D->bt_kcol.data();
  for (int k = 0; k < 3; ++k) {
    d.b[k] = D->b[k].data();
    d.bt[k] = D->bt[k].data();
  }
  d.u_begin = D->u_begin.data();
  d.u_global = D->u_global.data();
  d.p_begin = D->p_begin.data();
  d.p_global = D->p_global.data();
  d.parent = d.n_children ? D->parent.data() : nullptr;
  d.child = d.n_children ? D->child.data() : nullptr;
  d.qq_u = d.n_children ? D->qq_u.data() : nullptr;
  d.qq_p = d.n_children ? D->qq_p.data() : nullptr;
  d.ptr_colors = d.n_colors ? D->ptr_colors.data() : nullptr;
  d.cell_reorder = d.n_colors ? D->cell_reorder.data() : nullptr;
  *out = d;
  *handle = D.release();
  return PMG_OK;
}

int pmg_level_read_free(void *handle) {
  delete static_cast<Dump *>(handle);
  return PMG_OK;
}

int pmg_level_write(const char *dir, const pmg_level_desc *d) {
  if (!dir || !d) return PMG_ERR_INVALID_ARGUMENT;
  const std::string base(dir);
  const size_t nu = size_t(d->n_udof), np = size_t(d->n_pdof), nc = size_t(d->n_cells);
  const int nul = d->n_ulocal, npl = d->n_plocal;
  // BeginIndex with its closing entry; GlobalNumbers in cell order
  std::vector<int> ub(nc + 1), pb(nc + 1), ug, pg;
  for (size_t c = 0; c < nc; ++c) {
    ub[c] = int(ug.size());
    for (int j = 0; j < nul; ++j) ug.push_back(d->u_global[d->u_begin[c] + j]);
    pb[c] = int(pg.size());
    for (int j = 0; j < npl; ++j) pg.push_back(d->p_global[d->p_begin[c] + j]);
  }
  ub[nc] = int(ug.size());
  pb[nc] = int(pg.size());
  bool ok = write_raw(base + "/a_rowptr.i32", d->a_rowptr, nu + 1) && write_raw(base + "/a_kcol.i32", d->a_kcol, size_t(d->a_nnz)) &&
            write_raw(base + "/b_rowptr.i32", d->b_rowptr, np + 1) && write_raw(base + "/b_kcol.i32", d->b_kcol, size_t(d->b_nnz)) &&
            write_raw(base + "/bt_rowptr.i32", d->bt_rowptr, nu + 1) &&
            write_raw(base + "/bt_kcol.i32", d->bt_kcol, size_t(d->bt_nnz));
  for (int k = 0; ok && k < 9; ++k) ok = write_raw(base + "/" + kA[k] + ".f64", d->a[k], size_t(d->a_nnz));
  for (int k = 0; ok && k < 3; ++k)
    ok = write_raw(base + "/" + kB[k] + ".f64", d->b[k], size_t(d->b_nnz)) &&
         write_raw(base + "/" + kBT[k] + ".f64", d->bt[k], size_t(d->bt_nnz));
  ok = ok && write_raw(base + "/u_begin.i32", ub.data(), ub.size()) && write_raw(base + "/u_global.i32", ug.data(), ug.size()) &&
       write_raw(base + "/p_begin.i32", pb.data(), pb.size()) && write_raw(base + "/p_global.i32", pg.data(), pg.size());
  const bool fine = d->parent && d->child && d->qq_u && d->qq_p && d->n_children > 0;
  if (ok && fine) {
    const size_t nch = size_t(d->n_children);
    ok = write_raw(base + "/parent.i32", d->parent, nc) && write_raw(base + "/child.i32", d->child, nc) &&
         write_raw(base + "/qq_u.f64", d->qq_u, nch * nul * nul) && write_raw(base + "/qq_p.f64", d->qq_p, nch * npl * npl);
  }
  const bool colors = d->ptr_colors && d->cell_reorder && d->n_colors > 0;
  if (ok && colors)
    ok = write_raw(base + "/ptr_colors.i32", d->ptr_colors, size_t(d->n_colors) + 1) &&
         write_raw(base + "/cell_reorder.i32", d->cell_reorder, nc);
  if (!ok) return PMG_ERR_INVALID_ARGUMENT;
  std::ofstream mf(base + "/manifest.txt", std::ios::trunc);
  if (!mf) return PMG_ERR_INVALID_ARGUMENT;
  mf << "n_udof " << d->n_udof << "\nn_active " << d->n_active << "\nhanging_bound " << d->hanging_bound
     << "\nn_dirichlet " << d->n_dirichlet << "\nn_pdof " << d->n_pdof << "\na_nnz " << d->a_nnz << "\nb_nnz "
     << d->b_nnz << "\nbt_nnz " << d->bt_nnz << "\nn_cells " << d->n_cells << "\nn_ulocal " << nul << "\nn_plocal "
     << npl << "\nn_children " << (fine ? d->n_children : 0) << "\nn_colors " << (colors ? d->n_colors : 0) << "\n";
  return mf ? PMG_OK : PMG_ERR_INVALID_ARGUMENT;
}

int pmg_vector_read(const char *path, double *x, long n) {
  if (!path || !x || n < 0) return PMG_ERR_INVALID_ARGUMENT;
  std::ifstream f(path, std::ios::binary);
  if (!f) return PMG_ERR_INVALID_ARGUMENT;
  f.read(reinterpret_cast<char *>(x), std::streamsize(sizeof(double) * n));
  return size_t(f.gcount()) == sizeof(double) * size_t(n) ? PMG_OK : PMG_ERR_INVALID_ARGUMENT;
}

int pmg_vector_write(const char *path, const double *x, long n) {
  if (!path || !x || n < 0) return PMG_ERR_INVALID_ARGUMENT;
  return write_raw(path, x, size_t(n)) ? PMG_OK : PMG_ERR_INVALID_ARGUMENT;
}

}  // extern "C"
