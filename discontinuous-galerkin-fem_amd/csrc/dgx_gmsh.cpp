// GMSH ASCII meshes (MSH 2.2 and 4.1) on the host: the data MeshCreator::extract_mesh_data
// takes from a GMSH model (S/utils/mesh_creation.cpp:212-330) -- the nodes in file order, the
// 2-D elements of the first 2-D element type (3-node triangles or 4-node quadrilaterals), and
// the edges of the named 1-D physical groups, which become the boundary tags -- followed by the
// face connectivity of dgx_mesh_connect_elements without periodic pairing.
#include <algorithm>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "dgx_ctx.hpp"

namespace {

struct GmshMesh {
    std::vector<double> xy;                           // [n][2]
    std::vector<int32_t> elems;                       // [K][nv]
    int nv = 0;                                       // vertices per element
    std::vector<std::pair<int, std::string>> names;  // 1-D physical groups (tag, name), by tag
    std::map<std::pair<int, int>, int> edge_group;    // sorted vertex pair -> physical tag
};

// The text between "$Name" and "$EndName".
std::string section(const std::string& text, const std::string& name, bool required) {
    const std::string open = "$" + name, close = "$End" + name;
    const size_t a = text.find(open + "\n") != std::string::npos ? text.find(open + "\n") : text.find(open + "\r\n");
    if (a == std::string::npos) {
        if (required) throw std::invalid_argument("GMSH file has no " + open + " section");
        return {};
    }
    const size_t b = text.find(close, a);
    if (b == std::string::npos) throw std::invalid_argument("GMSH file has no " + close);
    return text.substr(a + open.size(), b - a - open.size());
}

int nodes_of(int type) {  // GMSH element types: 1 line, 2 triangle, 3 quadrangle, 15 point
    switch (type) {
    case 1: return 2;
    case 2: return 3;
    case 3: return 4;
    case 15: return 1;
    default: return -1;
    }
}

void read_names(const std::string& text, GmshMesh& g) {
    std::istringstream in(section(text, "PhysicalNames", false));
    int n = 0;
    if (!(in >> n)) return;
    for (int k = 0; k < n; ++k) {
        int dim = 0, tag = 0;
        in >> dim >> tag;
        std::string rest;
        std::getline(in, rest);
        const size_t q0 = rest.find('"'), q1 = rest.rfind('"');
        if (q0 == std::string::npos || q1 <= q0) throw std::invalid_argument("malformed $PhysicalNames entry");
        if (dim == 1) g.names.push_back({tag, rest.substr(q0 + 1, q1 - q0 - 1)});
    }
    std::sort(g.names.begin(), g.names.end());
}

// MSH 2.2: "$Nodes" n, then "tag x y z"; "$Elements" n, then "tag type ntags tags... nodes..."
// with the physical tag first.
void read_v2(const std::string& text, GmshMesh& g) {
    std::map<long, int> node;
    {
        std::istringstream in(section(text, "Nodes", true));
        long n = 0;
        in >> n;
        for (long k = 0; k < n; ++k) {
            long tag;
            double x, y, z;
            if (!(in >> tag >> x >> y >> z)) throw std::invalid_argument("truncated $Nodes section");
            node[tag] = int(g.xy.size() / 2);
            g.xy.push_back(x);
            g.xy.push_back(y);
        }
    }
    std::istringstream in(section(text, "Elements", true));
    long n = 0;
    in >> n;
    int type2d = 0;
    for (long k = 0; k < n; ++k) {
        long tag;
        int type, ntags;
        if (!(in >> tag >> type >> ntags)) throw std::invalid_argument("truncated $Elements section");
        std::vector<long> tags(ntags);
        for (auto& t : tags) in >> t;
        const int nn = nodes_of(type);
        if (nn < 0) throw std::invalid_argument("unsupported GMSH element type " + std::to_string(type));
        std::vector<int> v(nn);
        for (auto& x : v) {
            long t;
            in >> t;
            const auto it = node.find(t);
            if (it == node.end()) throw std::invalid_argument("element references a missing node");
            x = it->second;
        }
        if ((type == 2 || type == 3) && (type2d == 0 || type == type2d)) {
            type2d = type;
            g.elems.insert(g.elems.end(), v.begin(), v.end());
        } else if (type == 1 && ntags > 0) {
            g.edge_group[{std::min(v[0], v[1]), std::max(v[0], v[1])}] = int(tags[0]);
        }
    }
    g.nv = nodes_of(type2d);
}

// MSH 4.1: entities carry the physical tags; nodes and elements come in entity blocks.
void read_v4(const std::string& text, GmshMesh& g) {
    std::map<int, int> curve_group;  // curve entity -> its first physical tag
    {
        std::istringstream in(section(text, "Entities", true));
        long np, nc, ns, nvol;
        in >> np >> nc >> ns >> nvol;
        std::string line;
        std::getline(in, line);
        for (long k = 0; k < np; ++k) std::getline(in, line);  // points: not needed
        for (long k = 0; k < nc; ++k) {
            int tag, nphys;
            double b[6];
            in >> tag;
            for (double& x : b) in >> x;
            in >> nphys;
            std::vector<int> phys(nphys);
            for (int& x : phys) in >> x;
            if (nphys > 0) curve_group[tag] = phys[0];
            std::getline(in, line);  // bounding points
        }
    }
    std::map<long, int> node;
    {
        std::istringstream in(section(text, "Nodes", true));
        long nblocks, nnodes, mn, mx;
        in >> nblocks >> nnodes >> mn >> mx;
        for (long b = 0; b < nblocks; ++b) {
            int dim, etag, param;
            long cnt;
            if (!(in >> dim >> etag >> param >> cnt)) throw std::invalid_argument("truncated $Nodes section");
            std::vector<long> tags(cnt);
            for (auto& t : tags) in >> t;
            for (long k = 0; k < cnt; ++k) {
                double x, y, z, u;
                in >> x >> y >> z;
                if (param)
                    for (int d = 0; d < dim; ++d) in >> u;
                node[tags[k]] = int(g.xy.size() / 2);
                g.xy.push_back(x);
                g.xy.push_back(y);
            }
        }
    }
    std::istringstream in(section(text, "Elements", true));
    long nblocks, nel, mn, mx;
    in >> nblocks >> nel >> mn >> mx;
    int type2d = 0;
    for (long b = 0; b < nblocks; ++b) {
        int dim, etag, type;
        long cnt;
        if (!(in >> dim >> etag >> type >> cnt)) throw std::invalid_argument("truncated $Elements section");
        const int nn = nodes_of(type);
        if (nn < 0) throw std::invalid_argument("unsupported GMSH element type " + std::to_string(type));
        const bool take2d = dim == 2 && (type == 2 || type == 3) && (type2d == 0 || type == type2d);
        if (take2d) type2d = type;
        const auto cg = curve_group.find(etag);
        for (long k = 0; k < cnt; ++k) {
            long tag;
            in >> tag;
            std::vector<int> v(nn);
            for (auto& x : v) {
                long t;
                in >> t;
                const auto it = node.find(t);
                if (it == node.end()) throw std::invalid_argument("element references a missing node");
                x = it->second;
            }
            if (take2d) g.elems.insert(g.elems.end(), v.begin(), v.end());
            else if (dim == 1 && type == 1 && cg != curve_group.end())
                g.edge_group[{std::min(v[0], v[1]), std::max(v[0], v[1])}] = cg->second;
        }
    }
    g.nv = nodes_of(type2d);
}

GmshMesh read_gmsh(const char* path) {
    std::ifstream f(path);
    if (!f.is_open()) throw std::invalid_argument(std::string("Failed to open GMSH file: ") + path);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    std::istringstream fmt(section(text, "MeshFormat", true));
    double version = 0;
    int file_type = 0;
    fmt >> version >> file_type;
    if (file_type != 0) throw std::invalid_argument("binary GMSH files are not read; save as ASCII");
    GmshMesh g;
    read_names(text, g);
    if (version >= 4.0 && version < 5.0) read_v4(text, g);
    else if (version >= 2.0 && version < 3.0) read_v2(text, g);
    else throw std::invalid_argument("unsupported GMSH format version " + std::to_string(version));
    if (g.nv < 3 || g.elems.empty()) throw std::invalid_argument("No 2D elements found in GMSH mesh");
    return g;
}

}  // namespace

extern "C" dgx_status dgx_mesh_read_gmsh(const char* path, int32_t* n_vertices, int32_t* n_elements,
                                         int32_t* vertices_per_element, int32_t* n_tags, double* vertices,
                                         int32_t* elements, int32_t* face_neighbors, char* tag_names) {
    return guarded([&] {
        if (!path || !n_vertices || !n_elements || !vertices_per_element || !n_tags)
            throw std::invalid_argument("null argument");
        const GmshMesh g = read_gmsh(path);
        const int nvtx = int(g.xy.size() / 2), K = int(g.elems.size() / g.nv);
        *n_vertices = nvtx;
        *n_elements = K;
        *vertices_per_element = g.nv;
        *n_tags = int(g.names.size());
        if (!vertices && !elements && !face_neighbors && !tag_names) return;
        if (!vertices || !elements || !face_neighbors || !tag_names) throw std::invalid_argument("null output array");
        std::copy(g.xy.begin(), g.xy.end(), vertices);
        std::copy(g.elems.begin(), g.elems.end(), elements);
        for (size_t k = 0; k < g.names.size(); ++k) {
            std::memset(tag_names + 64 * k, 0, 64);
            std::strncpy(tag_names + 64 * k, g.names[k].second.c_str(), 63);
        }
        const dgx_status st = dgx_mesh_connect_elements(nvtx, vertices, K, elements, g.nv, 0, face_neighbors);
        if (st != DGX_OK) throw Fail(st, g_error);
        // boundary faces: the index of their physical group among the names, or -1
        std::map<int, int> index;
        for (size_t k = 0; k < g.names.size(); ++k) index[g.names[k].first] = int(k);
        for (int e = 0; e < K; ++e)
            for (int fc = 0; fc < g.nv; ++fc) {
                int32_t* nb = face_neighbors + (size_t(e) * g.nv + fc) * 2;
                if (nb[0] >= 0) continue;
                const int a = elements[e * g.nv + fc], b = elements[e * g.nv + (fc + 1) % g.nv];
                const auto it = g.edge_group.find({std::min(a, b), std::max(a, b)});
                nb[1] = -1;
                if (it != g.edge_group.end()) {
                    const auto ix = index.find(it->second);
                    if (ix != index.end()) nb[1] = ix->second;
                }
            }
    });
}
