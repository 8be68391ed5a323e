// C ABI over the host mesh layer (include/dkt.h).
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <exception>
#include <string>

#include "../../include/dkt.h"
#include "capi_common.h"
#include "mesh.h"

using namespace dkt;

namespace dkt {
thread_local std::string g_last_error;
}

extern "C" const char *dkt_last_error(void) { return dkt::g_last_error.c_str(); }

struct dkt_octs {
  MeshParams mp;
  std::vector<Oct> v;
};

struct dkt_da {
  MeshParams mp;
  std::vector<Oct> elems;
  NodeSet ns;
  // Built on first use, so a mesh pays only for the operator it is used with:
  // the reference traversal (reference-semantics plans, traversal queries)
  // and the conforming element records (conforming plans).
  mutable std::mutex mu;
  mutable std::unique_ptr<Traversal> ref, conf;
  Traversal &traversal() const {
    std::lock_guard<std::mutex> g(mu);
    if (!ref) ref = std::make_unique<Traversal>(flatten_traversal(mp, ns.nodes));
    return *ref;
  }
  const Traversal &conforming() const {
    std::lock_guard<std::mutex> g(mu);
    if (!conf) conf = std::make_unique<Traversal>(build_conforming(mp, elems, ns));
    return *conf;
  }
};

static int check_dims(int dim, int maxdepth) {
  if (dim < 2 || dim > 4) throw std::invalid_argument("dim must be 2, 3 or 4");
  if (maxdepth < 1 || maxdepth > 30) throw std::invalid_argument("maxdepth must be 1..30");
  return 0;
}

extern "C" int dkt_octs_create(int dim, int maxdepth, int curve, const uint32_t *cells, size_t n, dkt_octs **out) {
  return guarded([&] {
    check_dims(dim, maxdepth);
    if (!out || (n && !cells)) throw std::invalid_argument("null pointer");
    auto *o = new dkt_octs;
    o->mp.dim = dim;
    o->mp.maxd = maxdepth;
    o->mp.curve = curve;
    o->v.resize(n);
    const uint32_t top = 1u << maxdepth;
    for (size_t i = 0; i < n; i++) {
      Oct c{};
      for (int d = 0; d < dim; d++) {
        c.x[d] = cells[i * (dim + 1) + d];
        if (c.x[d] >= top) {
          delete o;
          throw std::invalid_argument("cell coordinate outside [0, 2^maxdepth)");
        }
      }
      c.lev = cells[i * (dim + 1) + dim];
      if ((int)c.lev > maxdepth) {
        delete o;
        throw std::invalid_argument("cell level above maxdepth");
      }
      const uint32_t mask = (c.lev == (uint32_t)maxdepth) ? 0u : ((1u << (maxdepth - c.lev)) - 1u);
      for (int d = 0; d < dim; d++) c.x[d] &= ~mask;
      o->v[i] = c;
    }
    *out = o;
  });
}

extern "C" int dkt_octs_destroy(dkt_octs *o) {
  delete o;
  return 0;
}
extern "C" size_t dkt_octs_size(const dkt_octs *o) { return o ? o->v.size() : 0; }
extern "C" int dkt_octs_copy(const dkt_octs *o, uint32_t *cells) {
  return guarded([&] {
    if (!o || !cells) throw std::invalid_argument("null pointer");
    const int dim = o->mp.dim;
    for (size_t i = 0; i < o->v.size(); i++) {
      for (int d = 0; d < dim; d++) cells[i * (dim + 1) + d] = o->v[i].x[d];
      cells[i * (dim + 1) + dim] = o->v[i].lev;
    }
  });
}

extern "C" int dkt_tree_sort(dkt_octs *o) {
  return guarded([&] {
    if (!o) throw std::invalid_argument("null pointer");
    sort_octants(o->mp, o->v, true);
  });
}

extern "C" int dkt_tree_construct(const dkt_octs *pts, size_t max_pts, dkt_octs **out) {
  return guarded([&] {
    if (!pts || !out) throw std::invalid_argument("null pointer");
    auto *t = new dkt_octs;
    t->mp = pts->mp;
    std::vector<Oct> tmp = pts->v;
    t->v = construct_tree(pts->mp, tmp, max_pts);
    *out = t;
  });
}

extern "C" int dkt_tree_balance(const dkt_octs *pts, size_t max_pts, dkt_octs **out) {
  return guarded([&] {
    if (!pts || !out) throw std::invalid_argument("null pointer");
    auto *t = new dkt_octs;
    t->mp = pts->mp;
    std::vector<Oct> tmp = pts->v;
    t->v = balance_tree(pts->mp, tmp, max_pts);
    *out = t;
  });
}

extern "C" int dkt_tree_sphere(int dim, int maxdepth, int curve, int base_level, int max_level, const double *center,
                               double radius, dkt_octs **out) {
  return guarded([&] {
    check_dims(dim, maxdepth);
    if (!center || !out) throw std::invalid_argument("null pointer");
    auto *t = new dkt_octs;
    t->mp.dim = dim;
    t->mp.maxd = maxdepth;
    t->mp.curve = curve;
    t->v = sphere_tree(t->mp, base_level, max_level, center, radius);
    *out = t;
  });
}

extern "C" int dkt_points_gauss(int dim, int maxdepth, int curve, size_t n, int ncomp, double sigma, uint64_t seed,
                                dkt_octs **out) {
  return guarded([&] {
    check_dims(dim, maxdepth);
    if (!out || ncomp < 1) throw std::invalid_argument("bad argument");
    auto *t = new dkt_octs;
    t->mp.dim = dim;
    t->mp.maxd = maxdepth;
    t->mp.curve = curve;
    t->v = gauss_points(t->mp, n, ncomp, sigma, seed);
    *out = t;
  });
}

extern "C" int dkt_keyed_uniform(uint64_t seed, int dim, const uint32_t *nodes, size_t n, double *x) {
  return guarded([&] {
    if (dim < 1 || dim > 4 || (n && (!nodes || !x))) throw std::invalid_argument("bad argument");
    for (size_t i = 0; i < n; i++) x[i] = keyed_uniform(seed, nodes + i * (dim + 1), dim);
  });
}

extern "C" int dkt_da_create(const dkt_octs *tree, int order, dkt_da **out) {
  return guarded([&] {
    if (!tree || !out) throw std::invalid_argument("null pointer");
    if (order < 1 || order > 2) throw std::invalid_argument("order must be 1 or 2");
    if (tree->v.empty()) throw std::invalid_argument("empty tree");
    auto *da = new dkt_da;
    try {
      da->mp = tree->mp;
      da->mp.order = order;
      da->elems = tree->v;
      for (const Oct &e : da->elems)
        if (order > 1 && (int)e.lev >= da->mp.maxd)
          throw std::invalid_argument("order 2 needs every element level below maxdepth");
      da->ns = enumerate_nodes(da->mp, da->elems);
    } catch (...) {
      delete da;
      throw;
    }
    *out = da;
  });
}

extern "C" int dkt_da_destroy(dkt_da *da) {
  delete da;
  return 0;
}

extern "C" size_t dkt_da_count(const dkt_da *da, int what) {
  if (!da) return 0;
  switch (what) {
    case 0: return da->elems.size();
    case 1: return da->ns.nodes.size();
    case 2: return da->ns.bdy.size();
    case 3: return da->traversal().leaves.size();
    case 4: return da->traversal().pslot.size() / (size_t)da->traversal().nPe;
    case 5: return (size_t)da->traversal().nPe;
    case 6: compute_stale_terms(da->mp, da->traversal()); return da->traversal().stale_needed;
    case 7: compute_stale_terms(da->mp, da->traversal()); return da->traversal().hc_node.size();
    case 8: return da->conforming().cf_node.size();
    case 9: return da->conforming().cf_off.size() - 1;
    default: return 0;
  }
}

static void copy_octs(const std::vector<Oct> &v, int dim, uint32_t *out) {
  for (size_t i = 0; i < v.size(); i++) {
    for (int d = 0; d < dim; d++) out[i * (dim + 1) + d] = v[i].x[d];
    out[i * (dim + 1) + dim] = v[i].lev;
  }
}

extern "C" int dkt_da_nodes(const dkt_da *da, uint32_t *out) {
  return guarded([&] {
    if (!da || !out) throw std::invalid_argument("null pointer");
    copy_octs(da->ns.nodes, da->mp.dim, out);
  });
}
extern "C" int dkt_da_elements(const dkt_da *da, uint32_t *out) {
  return guarded([&] {
    if (!da || !out) throw std::invalid_argument("null pointer");
    copy_octs(da->elems, da->mp.dim, out);
  });
}
extern "C" int dkt_da_boundary(const dkt_da *da, uint32_t *out) {
  return guarded([&] {
    if (!da || !out) throw std::invalid_argument("null pointer");
    std::memcpy(out, da->ns.bdy.data(), da->ns.bdy.size() * sizeof(uint32_t));
  });
}
extern "C" int dkt_da_traversal(const dkt_da *da, uint32_t *cells, uint8_t *child, int32_t *slot, int32_t *hang_of,
                                int32_t *pslot) {
  return guarded([&] {
    if (!da) throw std::invalid_argument("null pointer");
    const Traversal &t = da->traversal();
    if (cells) copy_octs(t.leaves, da->mp.dim, cells);
    if (child) std::memcpy(child, t.child.data(), t.child.size());
    if (slot) std::memcpy(slot, t.slot.data(), t.slot.size() * sizeof(int32_t));
    if (hang_of) std::memcpy(hang_of, t.hang_of.data(), t.hang_of.size() * sizeof(int32_t));
    if (pslot) std::memcpy(pslot, t.pslot.data(), t.pslot.size() * sizeof(int32_t));
  });
}

extern "C" int dkt_da_stale_terms(const dkt_da *da, int64_t *off, uint8_t *rank, int32_t *node, double *w) {
  return guarded([&] {
    if (!da) throw std::invalid_argument("null pointer");
    compute_stale_terms(da->mp, da->traversal());
    const Traversal &t = da->traversal();
    if (off) std::memcpy(off, t.hc_off.data(), t.hc_off.size() * sizeof(int64_t));
    if (rank) std::memcpy(rank, t.hc_rank.data(), t.hc_rank.size());
    if (node) std::memcpy(node, t.hc_node.data(), t.hc_node.size() * sizeof(int32_t));
    if (w) std::memcpy(w, t.hc_w.data(), t.hc_w.size() * sizeof(double));
  });
}

extern "C" int dkt_da_conforming_terms(const dkt_da *da, int64_t *off, uint8_t *rank, int32_t *node, double *w) {
  return guarded([&] {
    if (!da) throw std::invalid_argument("null pointer");
    const Traversal &t = da->conforming();
    if (off) std::memcpy(off, t.cf_off.data(), t.cf_off.size() * sizeof(int64_t));
    if (rank) std::memcpy(rank, t.cf_rank.data(), t.cf_rank.size());
    if (node) std::memcpy(node, t.cf_node.data(), t.cf_node.size() * sizeof(int32_t));
    if (w) std::memcpy(w, t.cf_w.data(), t.cf_w.size() * sizeof(double));
  });
}

extern "C" int dkt_da_write_vtu(const dkt_da *da, const char *path, const double *values, const char *name) {
  return guarded([&] {
    if (!da || !path) throw std::invalid_argument("null pointer");
    const int dim = da->mp.dim, p = da->mp.order, n1 = p + 1;
    if (dim != 2 && dim != 3) throw std::invalid_argument("VTU output supports dim 2 and 3");
    const Traversal &c = da->conforming();
    int npe = 1;
    for (int d = 0; d < dim; d++) npe *= n1;
    const int ncorner = 1 << dim;
    // VTK corner order (quad: 0 1 3 2; hexahedron: 0 1 3 2 4 5 7 6) of the
    // lexicographic corners, and the element rank of each corner.
    const int order2[4] = {0, 1, 3, 2}, order3[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    const int *vtk_order = dim == 2 ? order2 : order3;
    int corner_rank[8];
    for (int k = 0; k < ncorner; k++) {
      int r = 0, stride = 1;
      for (int d = 0; d < dim; d++, stride *= n1) r += ((k >> d) & 1) * p * stride;
      corner_rank[k] = r;
    }
    FILE *f = std::fopen(path, "w");
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    const size_t E = da->elems.size();
    const double scale = 1.0 / (double)(1u << da->mp.maxd);
    std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
    std::fprintf(f, "  <UnstructuredGrid>\n    <Piece NumberOfPoints=\"%zu\" NumberOfCells=\"%zu\">\n", E * ncorner, E);
    std::fprintf(f, "      <Points>\n        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n");
    for (size_t e = 0; e < E; e++) {
      const Oct &o = da->elems[e];
      const double len = (double)(1u << (da->mp.maxd - o.lev)) * scale;
      for (int v = 0; v < ncorner; v++) {
        const int k = vtk_order[v];
        double x[3] = {0, 0, 0};
        for (int d = 0; d < dim; d++) x[d] = o.x[d] * scale + ((k >> d) & 1) * len;
        std::fprintf(f, "%.17g %.17g %.17g\n", x[0], x[1], x[2]);
      }
    }
    std::fprintf(f, "        </DataArray>\n      </Points>\n      <Cells>\n");
    std::fprintf(f, "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n");
    for (size_t i = 0; i < E * ncorner; i++) std::fprintf(f, "%zu\n", i);
    std::fprintf(f, "        </DataArray>\n        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n");
    for (size_t e = 1; e <= E; e++) std::fprintf(f, "%zu\n", e * ncorner);
    std::fprintf(f, "        </DataArray>\n        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
    for (size_t e = 0; e < E; e++) std::fprintf(f, "%d\n", dim == 2 ? 9 : 12);
    std::fprintf(f, "        </DataArray>\n      </Cells>\n      <CellData Scalars=\"level\">\n");
    std::fprintf(f, "        <DataArray type=\"Int32\" Name=\"level\" format=\"ascii\">\n");
    for (size_t e = 0; e < E; e++) std::fprintf(f, "%u\n", da->elems[e].lev);
    std::fprintf(f, "        </DataArray>\n      </CellData>\n");
    if (values) {
      const char *nm = name && *name ? name : "u";
      std::fprintf(f, "      <PointData Scalars=\"%s\">\n        <DataArray type=\"Float64\" Name=\"%s\" format=\"ascii\">\n",
                   nm, nm);
      for (size_t e = 0; e < E; e++) {
        const int h = c.hang_of[e];
        for (int v = 0; v < ncorner; v++) {
          const int r = corner_rank[vtk_order[v]];
          const int32_t node = c.slot[e * npe + r];
          double val = 0.0;
          if (node >= 0) {
            val = values[node];
          } else {
            for (int64_t k = c.cf_off[h]; k < c.cf_off[h + 1]; k++)
              if (c.cf_rank[k] == r) val += c.cf_w[k] * values[c.cf_node[k]];
          }
          std::fprintf(f, "%.17g\n", val);
        }
      }
      std::fprintf(f, "        </DataArray>\n      </PointData>\n");
    }
    std::fprintf(f, "    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n");
    if (std::fclose(f) != 0) throw std::runtime_error(std::string("write failed: ") + path);
  });
}

// Internal accessors for the device plan (plan.cpp).
namespace dkt {
const MeshParams &da_params(const dkt_da *da) { return da->mp; }
const NodeSet &da_nodes(const dkt_da *da) { return da->ns; }
const Traversal &da_traversal(const dkt_da *da) { return da->traversal(); }
Traversal &da_traversal_mut(const dkt_da *da) { return da->traversal(); }
Traversal &da_conforming_mut(const dkt_da *da) { return const_cast<Traversal &>(da->conforming()); }
}  // namespace dkt
