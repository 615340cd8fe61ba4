// 3DGS PLY scene files and scene normalisation, on host slabs.
//
//   ds_ply_info / ds_ply_load   PLYLoader::load   (src/io/ply_io.cpp:15-160)
//   ds_ply_save                 PLYSaver::save    (src/io/ply_io.cpp:164-227)
//   ds_normalize_scene          SplatUtils::normalizeScene (src/utils/splat_utils.h:53-93)
//
// Format: binary little-endian, one `vertex` element of float properties
// addressed by name (x y z, scale_0..2, rot_0..3 = w x y z, f_dc_0..2,
// f_rest_* channel-first: all bands of R, then G, then B, opacity; others
// such as normals are skipped). Files store the OpenCV convention (y down,
// z forward); slabs use OpenGL (y up, z backward): y, z and the quaternion's
// y, z change sign on load and on save. The SH degree follows the f_rest
// count as in the reference (>= 45: 3, >= 24: 2, >= 9: 1, else 0). Unlike the
// reference, which would misread them silently, ASCII files and non-float
// vertex properties are rejected.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../../include/ds_api.h"
#include "../ds_internal.hpp"

namespace {

struct PlyHeader {
    int64_t count = 0;
    std::vector<std::string> props;  // float properties in file order
    std::streamoff data_offset = 0;
    int sh_degree = 0;
    int bands = 0;  // per channel
};

int bands_of_degree(int d) { return d == 3 ? 15 : d == 2 ? 8 : d == 1 ? 3 : 0; }

ds_status read_header(std::ifstream &f, const std::string &path, PlyHeader &h) {
    std::string line;
    if (!std::getline(f, line)) return ds::set_error(DS_ERR_INVALID_ARG, "empty file: " + path);
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line != "ply") return ds::set_error(DS_ERR_INVALID_ARG, "not a PLY file: " + path);
    bool in_vertex = false, done = false, binary_le = false;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::istringstream ss(line);
        std::string w0, w1, w2;
        ss >> w0 >> w1 >> w2;
        if (w0 == "format") {
            binary_le = (w1 == "binary_little_endian");
        } else if (w0 == "element") {
            in_vertex = (w1 == "vertex");
            if (in_vertex) h.count = std::stoll(w2);
        } else if (w0 == "property" && in_vertex) {
            if (w1 != "float" && w1 != "float32")
                return ds::set_error(DS_ERR_INVALID_ARG,
                                     "PLY vertex property '" + line + "' is not float: " + path);
            h.props.push_back(w2);
        } else if (w0 == "end_header") {
            done = true;
            break;
        }
    }
    if (!done) return ds::set_error(DS_ERR_INVALID_ARG, "malformed PLY, no end_header: " + path);
    if (!binary_le)
        return ds::set_error(DS_ERR_INVALID_ARG, "PLY is not binary_little_endian: " + path);
    if (h.count <= 0) return ds::set_error(DS_ERR_INVALID_ARG, "PLY has no vertices: " + path);
    h.data_offset = f.tellg();
    int rest = 0;
    for (const auto &p : h.props)
        if (p.rfind("f_rest_", 0) == 0) ++rest;
    const int per_channel = rest / 3;
    h.sh_degree = per_channel >= 15 ? 3 : per_channel >= 8 ? 2 : per_channel >= 3 ? 1 : 0;
    h.bands = bands_of_degree(h.sh_degree);
    return DS_OK;
}

ds_status open_ply(const char *path, std::ifstream &f, PlyHeader &h) {
    DS_REQUIRE(path, "NULL path");
    const std::string p(path);
    const size_t dot = p.find_last_of('.');
    std::string ext = dot == std::string::npos ? "" : p.substr(dot + 1);
    for (auto &c : ext) c = (char)std::tolower((unsigned char)c);
    DS_REQUIRE(ext == "ply", "unsupported file extension '" + ext + "': " + p);
    f.open(p, std::ios::binary);
    DS_REQUIRE(f.is_open(), "cannot open " + p);
    return read_header(f, p, h);
}

}  // namespace

extern "C" {

ds_status ds_ply_info(const char *path, int64_t *count, int32_t *sh_degree) {
    ds::clear_error();
    std::ifstream f;
    PlyHeader h;
    DS_TRY(open_ply(path, f, h));
    if (count) *count = h.count;
    if (sh_degree) *sh_degree = h.sh_degree;
    return DS_OK;
}

ds_status ds_ply_load(const char *path, float *host_slab, int64_t count, int32_t sh_degree) {
    ds::clear_error();
    DS_REQUIRE(host_slab, "NULL slab");
    std::ifstream f;
    PlyHeader h;
    DS_TRY(open_ply(path, f, h));
    DS_REQUIRE(count == h.count && sh_degree == h.sh_degree,
               "count / sh_degree differ from the file's (use ds_ply_info)");
    std::unordered_map<std::string, int> idx;
    for (int i = 0; i < (int)h.props.size(); ++i) idx[h.props[(size_t)i]] = i;
    const char *need[] = {"x",      "y",      "z",      "scale_0", "scale_1",
                          "scale_2", "rot_0", "rot_1",  "rot_2",   "rot_3",
                          "f_dc_0", "f_dc_1", "f_dc_2", "opacity"};
    int col[14];
    for (int k = 0; k < 14; ++k) {
        auto it = idx.find(need[k]);
        DS_REQUIRE(it != idx.end(), std::string("PLY is missing property ") + need[k]);
        col[k] = it->second;
    }
    const int nb = h.bands;
    std::vector<int> rest_col((size_t)3 * nb, -1);  // channel-first file order
    for (int b = 0; b < 3 * nb; ++b) {
        auto it = idx.find("f_rest_" + std::to_string(b));
        if (it != idx.end()) rest_col[(size_t)b] = it->second;
    }
    const int64_t n = h.count;
    const size_t np = h.props.size();
    std::vector<float> row(np);
    float *s = host_slab;
    float *rest = s + 13 * n;  // sh_rest_r, sh_rest_g, sh_rest_b: [band * n + i] each
    float *logit = rest + (int64_t)3 * nb * n;
    for (int64_t i = 0; i < n; ++i) {
        f.read(reinterpret_cast<char *>(row.data()), (std::streamsize)(np * sizeof(float)));
        DS_REQUIRE(f.gcount() == (std::streamsize)(np * sizeof(float)),
                   "PLY ends early at vertex " + std::to_string(i));
        // slab fields: pos xyz, scale xyz, rot wxyz, dc rgb
        s[0 * n + i] = row[(size_t)col[0]];
        s[1 * n + i] = -row[(size_t)col[1]];  // OpenCV -> OpenGL
        s[2 * n + i] = -row[(size_t)col[2]];
        for (int k = 0; k < 3; ++k) s[(3 + k) * n + i] = row[(size_t)col[3 + k]];
        s[6 * n + i] = row[(size_t)col[6]];
        s[7 * n + i] = row[(size_t)col[7]];
        s[8 * n + i] = -row[(size_t)col[8]];
        s[9 * n + i] = -row[(size_t)col[9]];
        for (int k = 0; k < 3; ++k) s[(10 + k) * n + i] = row[(size_t)col[10 + k]];
        for (int c = 0; c < 3; ++c)
            for (int b = 0; b < nb; ++b) {
                const int fc = rest_col[(size_t)(c * nb + b)];
                rest[((int64_t)c * nb + b) * n + i] = fc >= 0 ? row[(size_t)fc] : 0.f;
            }
        logit[i] = row[(size_t)col[13]];
    }
    return DS_OK;
}

ds_status ds_ply_save(const char *path, const float *host_slab, int64_t count,
                      int32_t sh_degree) {
    ds::clear_error();
    DS_REQUIRE(path && host_slab, "NULL argument");
    DS_REQUIRE(count > 0, "count must be positive");
    DS_REQUIRE(sh_degree >= 0 && sh_degree <= 3, "sh_degree must be in [0,3]");
    std::ofstream f(path, std::ios::binary);
    DS_REQUIRE(f.is_open(), std::string("cannot open for writing: ") + path);
    const int nb = bands_of_degree(sh_degree);
    const int64_t n = count;
    f << "ply\nformat binary_little_endian 1.0\nelement vertex " << n << "\n";
    auto prop = [&](const std::string &name) { f << "property float " << name << "\n"; };
    prop("x");
    prop("y");
    prop("z");
    prop("f_dc_0");
    prop("f_dc_1");
    prop("f_dc_2");
    for (int k = 0; k < 3 * nb; ++k) prop("f_rest_" + std::to_string(k));
    prop("opacity");
    prop("scale_0");
    prop("scale_1");
    prop("scale_2");
    prop("rot_0");
    prop("rot_1");
    prop("rot_2");
    prop("rot_3");
    f << "end_header\n";
    const float *s = host_slab;
    const float *rest = s + 13 * n;
    const float *logit = rest + (int64_t)3 * nb * n;
    std::vector<float> row((size_t)(14 + 3 * nb));
    for (int64_t i = 0; i < n; ++i) {
        size_t o = 0;
        row[o++] = s[0 * n + i];
        row[o++] = -s[1 * n + i];  // OpenGL -> OpenCV
        row[o++] = -s[2 * n + i];
        for (int k = 0; k < 3; ++k) row[o++] = s[(10 + k) * n + i];
        for (int c = 0; c < 3; ++c)
            for (int b = 0; b < nb; ++b) row[o++] = rest[((int64_t)c * nb + b) * n + i];
        row[o++] = logit[i];
        for (int k = 0; k < 3; ++k) row[o++] = s[(3 + k) * n + i];
        row[o++] = s[6 * n + i];
        row[o++] = s[7 * n + i];
        row[o++] = -s[8 * n + i];
        row[o++] = -s[9 * n + i];
        f.write(reinterpret_cast<const char *>(row.data()), (std::streamsize)(o * sizeof(float)));
    }
    DS_REQUIRE(f.good(), std::string("write failed: ") + path);
    return DS_OK;
}

// Sequential float arithmetic in the reference's order, so the result is
// bit-identical to SplatUtils::normalizeScene on the same splats.
ds_status ds_normalize_scene(float *host_slab, int64_t count, float scene_scale) {
    ds::clear_error();
    DS_REQUIRE(host_slab, "NULL slab");
    DS_REQUIRE(count > 0, "count must be positive");
    const int64_t n = count;
    float *px = host_slab, *py = host_slab + n, *pz = host_slab + 2 * n;
    float *sx = host_slab + 3 * n, *sy = host_slab + 4 * n, *sz = host_slab + 5 * n;
    float cx = 0.f, cy = 0.f, cz = 0.f;
    for (int64_t i = 0; i < n; ++i) {
        cx += px[i];
        cy += py[i];
        cz += pz[i];
    }
    const float inv = 1.f / (float)n;
    cx *= inv;
    cy *= inv;
    cz *= inv;
    float max_ext = 0.f;
    for (int64_t i = 0; i < n; ++i) {
        px[i] -= cx;
        py[i] -= cy;
        pz[i] -= cz;
        max_ext = std::max(max_ext, std::fabs(px[i]));
        max_ext = std::max(max_ext, std::fabs(py[i]));
        max_ext = std::max(max_ext, std::fabs(pz[i]));
    }
    if (max_ext > 0.f) {
        const float lin = scene_scale / max_ext;
        const float lg = std::log(lin);
        for (int64_t i = 0; i < n; ++i) {
            px[i] *= lin;
            py[i] *= lin;
            pz[i] *= lin;
            sx[i] += lg;
            sy[i] += lg;
            sz[i] += lg;
        }
    }
    return DS_OK;
}

}  // extern "C"
