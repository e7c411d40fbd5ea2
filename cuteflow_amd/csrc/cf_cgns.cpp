// CGNS mesh and solution files through the HDF5 C API.
//
// The CGNS library stores each CGNS node as an HDF5 group carrying string
// attributes "name" (33 bytes), "label" (33) and "type" (3: MT, C1, I4, I8,
// R4, R8) and an int attribute "flags"; a node with data holds it in a
// dataset named " data", with the CGNS dimensions in reverse order.  The
// root group carries "name"/"label"/"type" and the datasets " format" and
// " hdf5version".  Children are listed in link creation order, which the
// library tracks on every group.  This file reads and writes exactly the
// nodes the reference's mesh reader (lec_mail_cgns, maillage.cuf:398-778),
// mesh converter (cute_to_cgns.cpp:118-160), Manning writer
// (gen_manning.py:71-74) and FlowSolution export
// (pre_post_traitement.cuf:209-405) touch.

#include "../../include/cuteflow_cgns.h"

#include <hdf5.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// HDF5 prints its error stack on every failed call unless told otherwise;
// failures here are reported through the status and cf_cgns_error().
struct QuietH5 {
    H5E_auto2_t fn = nullptr;
    void *data = nullptr;
    QuietH5() {
        H5Eget_auto2(H5E_DEFAULT, &fn, &data);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~QuietH5() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

// Owned HDF5 identifier.
class Hid {
  public:
    Hid() = default;
    Hid(hid_t id, herr_t (*close)(hid_t)) : id_(id), close_(close) {}
    Hid(Hid &&o) noexcept : id_(o.id_), close_(o.close_) { o.id_ = -1; }
    Hid &operator=(Hid &&o) noexcept {
        if (this != &o) {
            reset();
            id_ = o.id_;
            close_ = o.close_;
            o.id_ = -1;
        }
        return *this;
    }
    Hid(const Hid &) = delete;
    Hid &operator=(const Hid &) = delete;
    ~Hid() { reset(); }
    void reset() {
        if (id_ >= 0 && close_) close_(id_);
        id_ = -1;
    }
    bool ok() const { return id_ >= 0; }
    operator hid_t() const { return id_; }

  private:
    hid_t id_ = -1;
    herr_t (*close_)(hid_t) = nullptr;
};

Hid open_group(hid_t loc, const std::string &name) {
    if (H5Lexists(loc, name.c_str(), H5P_DEFAULT) <= 0) return Hid();
    return Hid(H5Gopen2(loc, name.c_str(), H5P_DEFAULT), H5Gclose);
}

std::string attr_str(hid_t g, const char *key) {
    if (H5Aexists(g, key) <= 0) return "";
    Hid a(H5Aopen(g, key, H5P_DEFAULT), H5Aclose);
    if (!a.ok()) return "";
    Hid t(H5Aget_type(a), H5Tclose);
    if (H5Tget_class(t) != H5T_STRING || H5Tis_variable_str(t) > 0) return "";
    const size_t n = H5Tget_size(t);
    std::vector<char> buf(n + 1, 0);
    if (H5Aread(a, t, buf.data()) < 0) return "";
    return std::string(buf.data());
}

// A child node: HDF5 link name, CGNS name and label.
struct Node {
    std::string link, name, label;
};

herr_t collect_link(hid_t, const char *name, const H5L_info_t *, void *op) {
    static_cast<std::vector<std::string> *>(op)->push_back(name);
    return 0;
}

// Child nodes of g in the CGNS library's order (creation order; name order
// for groups that do not track it), without the " data"-style datasets.
std::vector<Node> children(hid_t g, const char *label = nullptr) {
    std::vector<std::string> links;
    hsize_t idx = 0;
    if (H5Literate(g, H5_INDEX_CRT_ORDER, H5_ITER_INC, &idx, collect_link, &links) < 0) {
        links.clear();
        idx = 0;
        H5Literate(g, H5_INDEX_NAME, H5_ITER_INC, &idx, collect_link, &links);
    }
    std::vector<Node> out;
    for (const auto &l : links) {
        if (l.empty() || l[0] == ' ') continue;
        H5O_info_t oi;
        if (H5Oget_info_by_name(g, l.c_str(), &oi, H5P_DEFAULT) < 0 || oi.type != H5O_TYPE_GROUP) continue;
        Hid c = open_group(g, l);
        if (!c.ok()) continue;
        Node n{l, attr_str(c, "name"), attr_str(c, "label")};
        if (n.name.empty()) n.name = l;
        if (!label || n.label == label) out.push_back(n);
    }
    return out;
}

const Node *find_named(const std::vector<Node> &v, const std::string &name) {
    for (const auto &n : v)
        if (n.name == name) return &n;
    return nullptr;
}

template <typename T>
hid_t mem_type();
template <>
hid_t mem_type<double>() { return H5T_NATIVE_DOUBLE; }
template <>
hid_t mem_type<long long>() { return H5T_NATIVE_LLONG; }
template <>
hid_t mem_type<signed char>() { return H5T_NATIVE_SCHAR; }

// The node's " data" array, converted to T; dims in HDF5 (reversed CGNS)
// order.  False when the node has no data or it cannot be read.
template <typename T>
bool read_data(hid_t g, std::vector<T> &out, std::vector<hsize_t> *dims = nullptr) {
    if (H5Lexists(g, " data", H5P_DEFAULT) <= 0) return false;
    Hid d(H5Dopen2(g, " data", H5P_DEFAULT), H5Dclose);
    if (!d.ok()) return false;
    Hid s(H5Dget_space(d), H5Sclose);
    const int nd = H5Sget_simple_extent_ndims(s);
    if (nd < 0) return false;
    std::vector<hsize_t> dd((size_t)nd);
    if (nd > 0) H5Sget_simple_extent_dims(s, dd.data(), nullptr);
    const hssize_t n = H5Sget_simple_extent_npoints(s);
    if (n < 0) return false;
    out.assign((size_t)n, T());
    if (n > 0 && H5Dread(d, mem_type<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) < 0) return false;
    if (dims) *dims = dd;
    return true;
}

// A C1 node's data as a string, trailing blanks and NULs removed.
std::string read_c1(hid_t g) {
    std::vector<signed char> v;
    if (!read_data(g, v)) return "";
    std::string s(v.begin(), v.end());
    while (!s.empty() && (s.back() == ' ' || s.back() == '\0')) s.pop_back();
    return s;
}

std::string rtrim(std::string s) {
    while (!s.empty() && (s.back() == ' ' || s.back() == '\0')) s.pop_back();
    return s;
}

// ---- the zone the reader takes -------------------------------------------

struct Zone {
    Hid file, base, zone;
    int nzones = 0, index = 0;
    long long nnt = 0, nelt = 0;
};

int open_zone(const char *path, unsigned mode, Zone &z) {
    if (!path) return fail(CF_ERR_ARG, "null path");
    z.file = Hid(H5Fopen(path, mode, H5P_DEFAULT), H5Fclose);
    if (!z.file.ok()) return fail(CF_ERR_IO, "%s: cannot open as an HDF5 file", path);
    Hid root(H5Gopen2(z.file, "/", H5P_DEFAULT), H5Gclose);
    const auto bases = children(root, "CGNSBase_t");
    if (bases.empty()) return fail(CF_ERR_IO, "%s: no CGNSBase_t node", path);
    z.base = open_group(root, bases[0].link);  // base = 1 (maillage.cuf:447)
    const auto zones = children(z.base, "Zone_t");
    z.nzones = (int)zones.size();
    // maillage.cuf:453-470: the first zone whose Ordinal equals the rank (0)
    for (size_t i = 0; i < zones.size(); i++) {
        Hid g = open_group(z.base, zones[i].link);
        if (!g.ok()) continue;
        long long ord = 0;
        const auto ords = children(g, "Ordinal_t");
        if (!ords.empty()) {
            Hid o = open_group(g, ords[0].link);
            std::vector<long long> v;
            if (o.ok() && read_data(o, v) && !v.empty()) ord = v[0];
        }
        if (ord == 0) {
            z.zone = std::move(g);
            z.index = (int)i + 1;
            break;
        }
    }
    if (!z.zone.ok()) return fail(CF_ERR_IO, "%s: no zone with ordinal 0 among %d zones", path, z.nzones);
    const auto zt = children(z.zone, "ZoneType_t");
    if (!zt.empty()) {
        Hid t = open_group(z.zone, zt[0].link);
        if (t.ok() && read_c1(t) != "Unstructured")
            return fail(CF_ERR_IO, "%s: zone %d is not Unstructured", path, z.index);
    }
    std::vector<long long> sizes;
    if (!read_data(z.zone, sizes) || sizes.size() < 2)
        return fail(CF_ERR_IO, "%s: zone %d has no sizes", path, z.index);
    z.nnt = sizes[0];
    z.nelt = sizes[1];
    if (z.nnt < 3 || z.nelt < 1 || z.nnt > 0x7fffffffLL || z.nelt > 0x7fffffffLL)
        return fail(CF_ERR_MESH, "%s: zone sizes %lld vertices, %lld cells", path, z.nnt, z.nelt);
    return CF_OK;
}

// A BC's point set: PointList / PointRange (or the older ElementList /
// ElementRange) as stored.  cg_boco_info counts a range as 2 points and
// cg_boco_read returns its two ends, which the reference then takes as
// node ids; the same values are returned here.
bool bc_points(hid_t bc, std::vector<long long> &pts) {
    static const char *names[] = {"PointList", "PointRange", "ElementList", "ElementRange"};
    const auto kids = children(bc);
    for (const char *nm : names) {
        const Node *n = find_named(kids, nm);
        if (!n) continue;
        Hid p = open_group(bc, n->link);
        return p.ok() && read_data(p, pts);
    }
    pts.clear();
    return true;
}

struct Bc {
    std::string name;
    int kind;  // 9 BCInflow, 13 BCOutflow (the CGNS BCType_t values), 0 other
    std::vector<long long> pts;
};

int read_bcs(const char *path, Zone &z, std::vector<Bc> &bcs) {
    bcs.clear();
    const auto zbc = children(z.zone, "ZoneBC_t");
    if (zbc.empty()) return CF_OK;
    Hid g = open_group(z.zone, zbc[0].link);
    if (!g.ok()) return fail(CF_ERR_IO, "%s: cannot open ZoneBC", path);
    for (const auto &n : children(g, "BC_t")) {
        Hid b = open_group(g, n.link);
        if (!b.ok()) return fail(CF_ERR_IO, "%s: cannot open BC %s", path, n.name.c_str());
        Bc bc;
        bc.name = n.name;
        const std::string type = read_c1(b);
        bc.kind = type == "BCInflow" ? 9 : type == "BCOutflow" ? 13 : 0;
        if (bc.kind && !bc_points(b, bc.pts))
            return fail(CF_ERR_IO, "%s: cannot read the points of BC %s", path, n.name.c_str());
        bcs.push_back(std::move(bc));
    }
    return CF_OK;
}

struct UserData {
    bool manning = false, partitioned = false;
    std::string manning_link;
};

UserData user_data(hid_t zone) {
    UserData u;
    for (const auto &n : children(zone, "UserDefinedData_t")) {
        if (n.name == "Manning") {
            u.manning = true;
            u.manning_link = n.link;
        } else if (n.name == "ElemsToSend" || n.name == "ElemsToRecv") {
            u.partitioned = true;
        }
    }
    return u;
}

// Fortran string comparison of a trimmed, left-adjusted BC name with a
// blank-padded name field: equal up to trailing blanks (and NUL padding).
bool name_matches(const std::string &bc, const char *field, int len) {
    std::string f(field, field + len);
    const size_t nul = f.find('\0');
    if (nul != std::string::npos) f.resize(nul);
    return rtrim(f) == bc;
}

std::string adjustl_trim(const std::string &s) {
    size_t b = 0;
    while (b < s.size() && s[b] == ' ') b++;
    return rtrim(s.substr(b));
}

// ---- writing ---------------------------------------------------------------

bool set_str_attr(hid_t g, const char *key, const std::string &val, size_t size) {
    Hid t(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(t, size);
    H5Tset_strpad(t, H5T_STR_NULLTERM);
    Hid s(H5Screate(H5S_SCALAR), H5Sclose);
    Hid a(H5Acreate2(g, key, t, s, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!a.ok()) return false;
    std::vector<char> buf(size, 0);
    std::memcpy(buf.data(), val.data(), std::min(val.size(), size - 1));
    return H5Awrite(a, t, buf.data()) >= 0;
}

Hid group_plist() {
    Hid p(H5Pcreate(H5P_GROUP_CREATE), H5Pclose);
    H5Pset_link_creation_order(p, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
    return p;
}

// New CGNS node under parent (an existing child of that name is removed).
Hid new_node(hid_t parent, const std::string &name, const char *label, const char *type) {
    if (name.empty() || name.size() > 32) return Hid();
    if (H5Lexists(parent, name.c_str(), H5P_DEFAULT) > 0 && H5Ldelete(parent, name.c_str(), H5P_DEFAULT) < 0)
        return Hid();
    Hid gcpl = group_plist();
    Hid g(H5Gcreate2(parent, name.c_str(), H5P_DEFAULT, gcpl, H5P_DEFAULT), H5Gclose);
    if (!g.ok()) return g;
    const int flags = 0;
    const hsize_t one = 1;
    Hid s(H5Screate_simple(1, &one, nullptr), H5Sclose);
    Hid a(H5Acreate2(g, "flags", H5T_STD_I32LE, s, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!a.ok() || H5Awrite(a, H5T_NATIVE_INT, &flags) < 0 || !set_str_attr(g, "label", label, 33) ||
        !set_str_attr(g, "name", name, 33) || !set_str_attr(g, "type", type, 3))
        return Hid();
    return g;
}

// " data" of a node: CGNS dims (fastest first) are stored reversed.
bool write_raw(hid_t g, hid_t file_type, hid_t mem, const void *v, std::vector<hsize_t> cgns_dims) {
    std::vector<hsize_t> dims(cgns_dims.rbegin(), cgns_dims.rend());
    Hid s(H5Screate_simple((int)dims.size(), dims.data(), nullptr), H5Sclose);
    Hid d(H5Dcreate2(g, " data", file_type, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    return d.ok() && H5Dwrite(d, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) >= 0;
}

bool write_i4(hid_t g, const std::vector<int> &v, std::vector<hsize_t> dims = {}) {
    if (dims.empty()) dims = {(hsize_t)v.size()};
    return write_raw(g, H5T_STD_I32LE, H5T_NATIVE_INT, v.data(), dims);
}

bool write_r8(hid_t g, const double *v, hsize_t n) {
    return write_raw(g, H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, v, {n});
}

bool write_c1(hid_t g, const std::string &s, std::vector<hsize_t> dims = {}) {
    if (dims.empty()) dims = {(hsize_t)s.size()};
    return write_raw(g, H5T_STD_I8LE, H5T_NATIVE_SCHAR, s.data(), dims);
}

// Node with data in one call; false on any failure.
bool node_r8(hid_t parent, const std::string &name, const char *label, const double *v, hsize_t n) {
    Hid g = new_node(parent, name, label, "R8");
    return g.ok() && write_r8(g, v, n);
}

bool node_c1(hid_t parent, const std::string &name, const char *label, const std::string &s) {
    Hid g = new_node(parent, name, label, "C1");
    return g.ok() && write_c1(g, s);
}

// BC node with a PointList of the given (1-based) points.
bool node_bc(hid_t zbc, const std::string &name, const char *type, const std::vector<int> &pts) {
    Hid b = new_node(zbc, name, "BC_t", "C1");
    if (!b.ok() || !write_c1(b, type)) return false;
    Hid p = new_node(b, "PointList", "IndexArray_t", "I4");
    return p.ok() && write_i4(p, pts, {1, (hsize_t)pts.size()});
}

// The BCs of one kind: a single BC named `base` when every number is 1 and
// there is one inlet (outlet), else one BC per number k, "<base> <k-1>".
bool write_bc_group(hid_t zbc, const char *base, const char *type, int nio, int n, const int *nd, const int *num) {
    if (n <= 0) return true;
    if (nio <= 1) return node_bc(zbc, base, type, std::vector<int>(nd, nd + n));
    for (int k = 1; k <= nio; k++) {
        std::vector<int> pts;
        for (int i = 0; i < n; i++)
            if ((num ? num[i] : 1) == k) pts.push_back(nd[i]);
        if (pts.empty()) continue;
        if (!node_bc(zbc, std::string(base) + " " + std::to_string(k - 1), type, pts)) return false;
    }
    return true;
}

std::string solution_name(int k) { return "Solution_" + std::to_string(k); }

}  // namespace

extern "C" {

const char *cf_cgns_error(void) { return g_err.c_str(); }

int cf_cgns_mesh_info(const char *path, cf_cgns_info *info) {
    QuietH5 quiet;
    if (!info) return fail(CF_ERR_ARG, "null info");
    std::memset(info, 0, sizeof *info);
    Zone z;
    int st = open_zone(path, H5F_ACC_RDONLY, z);
    if (st != CF_OK) return st;
    std::vector<Bc> bcs;
    if ((st = read_bcs(path, z, bcs)) != CF_OK) return st;
    info->nnt = (int)z.nnt;
    info->nelt = (int)z.nelt;
    info->nzones = z.nzones;
    info->zone = z.index;
    for (const auto &b : bcs) {
        if (b.pts.size() <= 1) continue;  // n_bcnodes > 1 (maillage.cuf:541)
        if (b.kind == 9) info->ndi += (int)b.pts.size();
        if (b.kind == 13) info->ndo += (int)b.pts.size();
    }
    const UserData u = user_data(z.zone);
    info->has_manning = u.manning;
    info->partitioned = u.partitioned;
    return CF_OK;
}

int cf_cgns_read_mesh(const char *path, int nombre_entrees, const char *inlet_name, int nombre_sorties,
                      const char *outlet_name, int name_len, int is_override_manning, int verbose,
                      double *coords, int *conn, double *manning, int *ndinput, int *numero_ndinput,
                      int *ndoutput, int *numero_ndoutput) {
    QuietH5 quiet;
    if (!coords || !conn) return fail(CF_ERR_ARG, "null coords or conn");
    if (nombre_entrees < 0 || nombre_sorties < 0 || (nombre_entrees && !inlet_name) ||
        (nombre_sorties && !outlet_name) || ((nombre_entrees || nombre_sorties) && name_len <= 0))
        return fail(CF_ERR_ARG, "inlet/outlet names");
    Zone z;
    int st = open_zone(path, H5F_ACC_RDONLY, z);
    if (st != CF_OK) return st;
    const UserData u = user_data(z.zone);
    if (u.partitioned)
        return fail(CF_ERR_MESH,
                    "%s: zone %d holds ElemsToSend/ElemsToRecv, i.e. it is one part of a mesh partitioned "
                    "for several ranks; its ghost cells need the halo exchange, which this build does not do",
                    path, z.index);
    const long long nnt = z.nnt, nelt = z.nelt;

    // coordinates (maillage.cuf:477-499)
    const auto gcs = children(z.zone, "GridCoordinates_t");
    const Node *gcn = find_named(gcs, "GridCoordinates");
    if (!gcn) return fail(CF_ERR_IO, "%s: no GridCoordinates", path);
    Hid gc = open_group(z.zone, gcn->link);
    const auto cnodes = children(gc, "DataArray_t");
    static const char *cname[3] = {"CoordinateX", "CoordinateY", "CoordinateZ"};
    for (int d = 0; d < 3; d++) {
        const Node *n = find_named(cnodes, cname[d]);
        std::vector<double> v;
        Hid c = n ? open_group(gc, n->link) : Hid();
        if (!c.ok() || !read_data(c, v)) return fail(CF_ERR_IO, "%s: cannot read %s", path, cname[d]);
        if ((long long)v.size() != nnt)
            return fail(CF_ERR_MESH, "%s: %s holds %zu values for %lld vertices", path, cname[d], v.size(), nnt);
        for (long long i = 0; i < nnt; i++) coords[3 * i + d] = v[(size_t)i];
    }

    // section 1 (maillage.cuf:503-521)
    const auto secs = children(z.zone, "Elements_t");
    if (secs.empty()) return fail(CF_ERR_IO, "%s: no Elements_t section", path);
    Hid sec = open_group(z.zone, secs[0].link);
    std::vector<long long> sdata, range, cv;
    if (!sec.ok() || !read_data(sec, sdata) || sdata.empty())
        return fail(CF_ERR_IO, "%s: cannot read section %s", path, secs[0].name.c_str());
    const auto skids = children(sec);
    const Node *rn = find_named(skids, "ElementRange");
    const Node *cn = find_named(skids, "ElementConnectivity");
    Hid rg = rn ? open_group(sec, rn->link) : Hid();
    Hid cg = cn ? open_group(sec, cn->link) : Hid();
    if (!rg.ok() || !read_data(rg, range) || range.size() != 2 || !cg.ok() || !read_data(cg, cv))
        return fail(CF_ERR_IO, "%s: cannot read the range/connectivity of section %s", path, secs[0].name.c_str());
    const long long count = range[1] - range[0] + 1;
    if (sdata[0] != 5 || count != nelt || (long long)cv.size() != 3 * nelt)
        return fail(CF_ERR_MESH,
                    "%s: section %s holds %lld elements of type %lld; one TRI_3 (5) section over the "
                    "zone's %lld cells is required",
                    path, secs[0].name.c_str(), count, sdata[0], nelt);
    for (size_t i = 0; i < cv.size(); i++) {
        if (cv[i] < 1 || cv[i] > nnt)
            return fail(CF_ERR_MESH, "%s: connectivity entry %zu = %lld outside 1..%lld", path, i, cv[i], nnt);
        conn[i] = (int)cv[i];
    }

    // boundary conditions (maillage.cuf:525-616)
    std::vector<Bc> bcs;
    if ((st = read_bcs(path, z, bcs)) != CF_OK) return st;
    int ndi = 0, ndo = 0;
    for (const auto &b : bcs) {
        if (b.pts.size() <= 1 || !b.kind) continue;
        const bool in = b.kind == 9;
        int *nd = in ? ndinput : ndoutput;
        int *num = in ? numero_ndinput : numero_ndoutput;
        if (!nd || !num) return fail(CF_ERR_ARG, "null %s node arrays", in ? "inflow" : "outflow");
        const int nio = in ? nombre_entrees : nombre_sorties;
        const char *names = in ? inlet_name : outlet_name;
        const std::string bname = adjustl_trim(b.name);
        int which = 0;
        for (int it = 0; it < nio; it++)
            if (name_matches(bname, names + (size_t)it * name_len, name_len)) which = it + 1;
        if (!which) {
            if (verbose)
                printf(" %s name : %s, not found in donnees.f file ! Setting it to 1\n", in ? "Inlet" : "Outlet",
                       bname.c_str());
            which = 1;
        }
        int &off = in ? ndi : ndo;
        for (long long p : b.pts) {
            nd[off] = (int)p;
            num[off] = which;
            off++;
        }
    }

    // Manning (maillage.cuf:725-736)
    if (manning) {
        for (long long i = 0; i < nelt; i++) manning[i] = 0.0;
        if (u.manning) {
            if (is_override_manning == 1) {
                if (verbose) printf(" MANNING DETECTED IN MESH FILE BUT WILL NOT BE USED\n");
            } else {
                if (verbose) printf(" MANNING DETECTED IN MESH FILE WILL BE USED\n");
                Hid m = open_group(z.zone, u.manning_link);
                const auto arrays = m.ok() ? children(m, "DataArray_t") : std::vector<Node>();
                std::vector<double> v;
                Hid a = arrays.empty() ? Hid() : open_group(m, arrays[0].link);
                if (!a.ok() || !read_data(a, v)) return fail(CF_ERR_IO, "%s: cannot read the Manning array", path);
                if ((long long)v.size() != nelt)
                    return fail(CF_ERR_MESH, "%s: Manning holds %zu values for %lld cells", path, v.size(), nelt);
                for (long long i = 0; i < nelt; i++) manning[i] = v[(size_t)i];
            }
        }
    }
    if (verbose) fflush(stdout);
    return CF_OK;
}

int cf_cgns_write_mesh(const char *path, int nnt, int nelt, const double *coords, const int *conn,
                       int nombre_entrees, int ndi, const int *ndinput, const int *numero_ndinput,
                       int nombre_sorties, int ndo, const int *ndoutput, const int *numero_ndoutput,
                       int ndw, const int *ndwall, const double *manning) {
    QuietH5 quiet;
    if (!path || nnt < 3 || nelt < 1 || !coords || !conn || ndi < 0 || ndo < 0 || ndw < 0 ||
        (ndi && !ndinput) || (ndo && !ndoutput) || (ndw && !ndwall))
        return fail(CF_ERR_ARG, "invalid mesh arguments");
    for (long long i = 0; i < 3LL * nelt; i++)
        if (conn[i] < 1 || conn[i] > nnt) return fail(CF_ERR_MESH, "connectivity entry %lld = %d", i, conn[i]);

    Hid fcpl(H5Pcreate(H5P_FILE_CREATE), H5Pclose);
    H5Pset_link_creation_order(fcpl, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
    Hid f(H5Fcreate(path, H5F_ACC_TRUNC, fcpl, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(CF_ERR_IO, "%s: cannot create", path);
    Hid root(H5Gopen2(f, "/", H5P_DEFAULT), H5Gclose);
    unsigned maj = 0, min = 0, rel = 0;
    H5get_libversion(&maj, &min, &rel);
    char ver[33] = {0};
    snprintf(ver, sizeof ver, "HDF5 Version %u.%u.%u", maj, min, rel);
    const char format[7] = {'N', 'A', 'T', 'I', 'V', 'E', 0};
    bool ok = set_str_attr(root, "label", "Root Node of HDF5 File", 33) &&
              set_str_attr(root, "name", "HDF5 MotherNode", 33) && set_str_attr(root, "type", "MT", 3);
    {
        const hsize_t n7 = 7, n33 = 33;
        Hid s7(H5Screate_simple(1, &n7, nullptr), H5Sclose), s33(H5Screate_simple(1, &n33, nullptr), H5Sclose);
        Hid d1(H5Dcreate2(root, " format", H5T_STD_I8LE, s7, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
        Hid d2(H5Dcreate2(root, " hdf5version", H5T_STD_I8LE, s33, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
               H5Dclose);
        ok = ok && d1.ok() && d2.ok() && H5Dwrite(d1, H5T_NATIVE_SCHAR, H5S_ALL, H5S_ALL, H5P_DEFAULT, format) >= 0 &&
             H5Dwrite(d2, H5T_NATIVE_SCHAR, H5S_ALL, H5S_ALL, H5P_DEFAULT, ver) >= 0;
    }
    {
        Hid v = new_node(root, "CGNSLibraryVersion", "CGNSLibraryVersion_t", "R4");
        const float lv = 4.0f;  // the value the reference's mesh files hold
        ok = ok && v.ok() && write_raw(v, H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, &lv, {1});
    }
    Hid base = new_node(root, "Base", "CGNSBase_t", "I4");
    ok = ok && base.ok() && write_i4(base, {2, 3});
    Hid zone = ok ? new_node(base, "Zone", "Zone_t", "I4") : Hid();
    ok = ok && zone.ok() && write_i4(zone, {nnt, nelt, 0}, {1, 3});
    ok = ok && node_c1(zone, "ZoneType", "ZoneType_t", "Unstructured");
    if (ok) {
        Hid gc = new_node(zone, "GridCoordinates", "GridCoordinates_t", "MT");
        ok = gc.ok();
        static const char *cname[3] = {"CoordinateX", "CoordinateY", "CoordinateZ"};
        std::vector<double> v((size_t)nnt);
        for (int d = 0; ok && d < 3; d++) {
            for (int i = 0; i < nnt; i++) v[i] = coords[3 * (size_t)i + d];
            ok = node_r8(gc, cname[d], "DataArray_t", v.data(), (hsize_t)nnt);
        }
    }
    if (ok) {
        Hid sec = new_node(zone, "Elements", "Elements_t", "I4");
        ok = sec.ok() && write_i4(sec, {5, 0});
        Hid r = ok ? new_node(sec, "ElementRange", "IndexRange_t", "I4") : Hid();
        ok = ok && r.ok() && write_i4(r, {1, nelt});
        Hid c = ok ? new_node(sec, "ElementConnectivity", "DataArray_t", "I4") : Hid();
        ok = ok && c.ok() && write_i4(c, std::vector<int>(conn, conn + 3 * (size_t)nelt));
    }
    if (ok && (ndi || ndo || ndw)) {
        Hid zbc = new_node(zone, "ZoneBC", "ZoneBC_t", "MT");
        ok = zbc.ok() &&
             write_bc_group(zbc, "Inflow nodes", "BCInflow", nombre_entrees, ndi, ndinput, numero_ndinput) &&
             write_bc_group(zbc, "Outflow nodes", "BCOutflow", nombre_sorties, ndo, ndoutput, numero_ndoutput) &&
             (ndw == 0 || node_bc(zbc, "Wall nodes", "BCWall", std::vector<int>(ndwall, ndwall + ndw)));
    }
    if (ok && manning) {
        Hid m = new_node(zone, "Manning", "UserDefinedData_t", "MT");
        ok = m.ok() && node_c1(m, "GridLocation", "GridLocation_t", "CellCenter") &&
             node_r8(m, "{DataArray}", "DataArray_t", manning, (hsize_t)nelt);
    }
    if (!ok) return fail(CF_ERR_IO, "%s: write failed", path);
    return H5Fflush(f, H5F_SCOPE_LOCAL) < 0 ? fail(CF_ERR_IO, "%s: flush failed", path) : CF_OK;
}

int cf_cgns_export_init(const char *path, int nsnap, const double *times) {
    QuietH5 quiet;
    if (nsnap < 1 || !times) return fail(CF_ERR_ARG, "nsnap %d", nsnap);
    Zone z;
    int st = open_zone(path, H5F_ACC_RDWR, z);
    if (st != CF_OK) return st;
    Hid bi = new_node(z.base, "TimeIterValues", "BaseIterativeData_t", "I4");
    bool ok = bi.ok() && write_i4(bi, {nsnap}) && node_r8(bi, "TimeValues", "DataArray_t", times, (hsize_t)nsnap);
    Hid zi = ok ? new_node(z.zone, "ZoneIterativeData", "ZoneIterativeData_t", "MT") : Hid();
    ok = ok && zi.ok();
    if (ok) {
        std::string names;
        for (int k = 0; k < nsnap; k++) {
            std::string s = solution_name(k);
            s.resize(32, ' ');
            names += s;
        }
        Hid p = new_node(zi, "FlowSolutionPointers", "DataArray_t", "C1");
        ok = p.ok() && write_c1(p, names, {32, (hsize_t)nsnap});
    }
    if (!ok) return fail(CF_ERR_IO, "%s: writing the iterative data failed", path);
    return CF_OK;
}

int cf_cgns_export(const char *path, int k, int nelt, const double *U, const double *zm, const double *manning,
                   double tolisec, double sol_z_offset) {
    QuietH5 quiet;
    if (k < 0 || nelt < 1 || !U || !zm) return fail(CF_ERR_ARG, "invalid export arguments");
    Zone z;
    int st = open_zone(path, H5F_ACC_RDWR, z);
    if (st != CF_OK) return st;
    if (z.nelt != nelt) return fail(CF_ERR_ARG, "%s: zone has %lld cells, state %d", path, z.nelt, nelt);
    // vdlg_visu (pre_post_traitement.cuf:344-366)
    const size_t n = (size_t)nelt;
    std::vector<double> wd(n), fs(n), mx(n), my(n), vx(n), vy(n), zero(n, 0.0);
    for (size_t i = 0; i < n; i++) {
        const double h = U[3 * i], hu = U[3 * i + 1], hv = U[3 * i + 2];
        wd[i] = h;
        if (h > tolisec) {
            mx[i] = hu;
            my[i] = hv;
            vx[i] = hu / h;
            vy[i] = hv / h;
        } else {
            mx[i] = my[i] = vx[i] = vy[i] = 0.0;
        }
        fs[i] = h + zm[i] - sol_z_offset;
    }
    Hid sol = new_node(z.zone, solution_name(k), "FlowSolution_t", "MT");
    bool ok = sol.ok() && node_c1(sol, "GridLocation", "GridLocation_t", "CellCenter") &&
              node_r8(sol, "Water Depth", "DataArray_t", wd.data(), n) &&
              node_r8(sol, "Bathymetry", "DataArray_t", zm, n) &&
              node_r8(sol, "Free surface elevation", "DataArray_t", fs.data(), n) &&
              (!manning || node_r8(sol, "Manning", "DataArray_t", manning, n)) &&
              node_r8(sol, "MomentumX", "DataArray_t", mx.data(), n) &&
              node_r8(sol, "MomentumY", "DataArray_t", my.data(), n) &&
              node_r8(sol, "MomentumZ", "DataArray_t", zero.data(), n) &&
              node_r8(sol, "VelocityX", "DataArray_t", vx.data(), n) &&
              node_r8(sol, "VelocityY", "DataArray_t", vy.data(), n) &&
              node_r8(sol, "VelocityZ", "DataArray_t", zero.data(), n);
    if (!ok) return fail(CF_ERR_IO, "%s: writing %s failed", path, solution_name(k).c_str());
    return CF_OK;
}

}  // extern "C"
