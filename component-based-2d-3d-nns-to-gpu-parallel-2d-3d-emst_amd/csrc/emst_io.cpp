// emst_io.cpp -- the reference's on-disk formats, host side, on top of the
// public C ABI (include/emst.h).
//
// Readers follow InputRW::readFile (emst_coalition/include/InputRW.h:44-623):
// the file name is cut at its first '.' or '_' (getPos, :44-58); the rest
// selects the reader (".points" 3D-image list, ".3d", ".2d", anything else
// TSPLIB) and the reader opens base + its own suffix.  Writers follow
// NeuralNetEMST::writeLinks (NeuralNetEMST.h:277-326), BufferLink's
// operator<< (BufferLink.h:266-279) and the Grid stream layout ("Width = W
// Height = H " header, one row of space-terminated values, blank line).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/emst.h"

namespace {

// InputRW::getPos: the first '.' or '_', whichever comes first -- searched
// in the file name only (the reference runs on bare names in the working
// directory; a directory part may hold either character)
size_t base_pos(const std::string& s)
{
    const size_t slash = s.find_last_of('/');
    const size_t from = slash == std::string::npos ? 0 : slash + 1;
    size_t pos = s.find('.', from);
    if (pos == std::string::npos) pos = s.size();
    const size_t t = s.find('_', from);
    if (t != std::string::npos && t < pos) pos = t;
    return pos;
}

// ostream << float with the default precision (6 significant digits, %g)
void put_g(std::string& o, double v)
{
    char b[64];
    std::snprintf(b, sizeof b, "%g", v);
    o += b;
}

// "Width = N Height = 1 \n", the way Grid's operator<< opens a map
std::string grid_header(int64_t n)
{
    return "Width = " + std::to_string(n) + " Height = 1 \n";
}

bool write_file(const std::string& path, const std::string& body)
{
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(body.data(), 1, body.size(), f) == body.size();
    return std::fclose(f) == 0 && ok;
}

// Bytes of the open file (its size bounds the counts its header may claim:
// every coordinate takes at least two characters).
long long file_bytes(std::ifstream& fi)
{
    const std::streampos here = fi.tellg();
    fi.seekg(0, std::ios::end);
    const long long n = (long long)fi.tellg();
    fi.seekg(here);
    return n < 0 ? 0 : n;
}

// readUniform (InputRW.h:516-623): first token N, then N points
int read_uniform(std::ifstream& fi, int dim, std::vector<float>& out)
{
    const long long bytes = file_bytes(fi);
    long long size = 0;
    if (!(fi >> size) || size < 0 || size > bytes / (2 * dim)) return EMST_E_IO;
    out.resize((size_t)size * dim);
    for (long long i = 0; i < size * dim; ++i)
        if (!(fi >> out[(size_t)i])) return EMST_E_IO;
    return EMST_OK;
}

// readTSP (InputRW.h:411-513): NAME on line 1, DIMENSION on line 4, two more
// header lines, then "index x y" per city
int read_tsp(std::ifstream& fi, std::vector<float>& xy)
{
    const long long bytes = file_bytes(fi);
    std::string line;
    if (!std::getline(fi, line) || line.find("NAME") == std::string::npos) return EMST_E_IO;
    for (int l = 1; l <= 3; ++l)
        if (!std::getline(fi, line)) return EMST_E_IO;
    const size_t d = line.find("DIMENSION");
    if (d == std::string::npos) return EMST_E_IO;
    char w1[256] = {0}, w2[256] = {0};
    long long size = 0;
    if (std::sscanf(line.c_str() + d, "%255s%255s%lld", w1, w2, &size) != 3 || size < 0) return EMST_E_IO;
    size = std::min(size, bytes / 6);       // an "index x y" row takes at least 6 characters
    for (int l = 1; l <= 2; ++l)
        if (!std::getline(fi, line)) return EMST_E_IO;
    // a file holding fewer cities than DIMENSION ends at its "EOF" line: the
    // cities present are kept (the reference would leave the rest of its
    // map at its initial value)
    xy.clear();
    for (long long i = 0; i < size; ++i) {
        long long num = 0;
        float x = 0.f, y = 0.f;
        if (!(fi >> num)) break;
        if (!(fi >> x >> y)) return EMST_E_IO;
        xy.push_back(x);
        xy.push_back(y);
    }
    return EMST_OK;
}

// read3DImage (InputRW.h:121-247): "Width = W Height = H", then W*H points of
// dim coordinates.  (The reference extracts the header words into an int,
// which fails on its own "Width = ..." header and reads no point; the header
// is parsed as written here.)
int read_points_list(std::ifstream& fi, int dim, std::vector<float>& out)
{
    const long long bytes = file_bytes(fi);
    std::string w, eq;
    long long width = 0, height = 0;
    if (!(fi >> w >> eq >> width) || !(fi >> w >> eq >> height) || width < 0 || height < 0) return EMST_E_IO;
    if (width > bytes || height > bytes) return EMST_E_IO;
    const long long size = width * height;
    if (size > bytes / (2 * dim)) return EMST_E_IO;
    out.resize((size_t)size * dim);
    for (long long i = 0; i < size; ++i)
        for (int k = 0; k < dim; ++k)
            if (!(fi >> out[(size_t)(i * dim + k)])) return EMST_E_IO;
    return EMST_OK;
}

}  // namespace

extern "C" {

static int read_points_impl(const char* path, int32_t dim, float* xyz, int64_t* n);

// The readers and writers never let an exception leave the C ABI: a failed
// allocation is EMST_E_OOM, anything else a file error.
int emst_read_points(const char* path, int32_t dim, float* xyz, int64_t* n)
{
    try {
        return read_points_impl(path, dim, xyz, n);
    } catch (const std::bad_alloc&) {
        return EMST_E_OOM;
    } catch (...) {
        return EMST_E_IO;
    }
}

static int read_points_impl(const char* path, int32_t dim, float* xyz, int64_t* n)
{
    if (!path || !n || (dim != 2 && dim != 3)) return EMST_E_ARG;
    const std::string file(path);
    const size_t pos = base_pos(file);
    const std::string base = file.substr(0, pos), sfx = file.substr(pos);
    std::vector<float> pts;      // dim coordinates per point
    int rc;
    if (sfx == ".points") {
        std::ifstream fi(base + ".points");
        if (!fi) return EMST_E_IO;
        rc = read_points_list(fi, dim, pts);
    } else if (sfx == ".3d") {
        if (dim != 3) return EMST_E_ARG;       // three coordinates per line
        std::ifstream fi(base + ".3d");
        if (!fi) return EMST_E_IO;
        rc = read_uniform(fi, 3, pts);
    } else if (sfx == ".2d") {
        std::ifstream fi(base + ".2d");
        if (!fi) return EMST_E_IO;
        std::vector<float> xy;
        rc = read_uniform(fi, 2, xy);
        if (rc == EMST_OK) {
            // a 3D solution reads the plane at z = 0 (read2DUniform, InputRW.h:91-95)
            const size_t m = xy.size() / 2;
            pts.assign(m * dim, 0.f);
            for (size_t i = 0; i < m; ++i)
                for (int k = 0; k < 2; ++k) pts[i * dim + k] = xy[2 * i + k];
        }
    } else {
        std::ifstream fi(base + ".tsp");
        if (!fi) return EMST_E_IO;
        std::vector<float> xy;
        rc = read_tsp(fi, xy);
        if (rc == EMST_OK) {
            const size_t m = xy.size() / 2;
            pts.assign(m * dim, 0.f);
            for (size_t i = 0; i < m; ++i)
                for (int k = 0; k < 2; ++k) pts[i * dim + k] = xy[2 * i + k];
        }
    }
    if (rc != EMST_OK) return rc;
    const int64_t m = (int64_t)(pts.size() / dim);
    if (!xyz) {
        *n = m;
        return EMST_OK;
    }
    if (*n < m) {
        *n = m;
        return EMST_E_ARG;
    }
    std::memcpy(xyz, pts.data(), sizeof(float) * pts.size());
    *n = m;
    return EMST_OK;
}

static int write_solution_impl(emst_ctx* ctx, const char* path);

int emst_write_solution(emst_ctx* ctx, const char* path)
{
    try {
        return write_solution_impl(ctx, path);
    } catch (const std::bad_alloc&) {
        return EMST_E_OOM;
    } catch (...) {
        return EMST_E_IO;
    }
}

static int write_solution_impl(emst_ctx* ctx, const char* path)
{
    if (!ctx || !path) return EMST_E_ARG;
    const std::string file(path);
    const std::string base = file.substr(0, base_pos(file));
    // the forest and the maps, from the context
    int64_t m = 0;
    int rc = emst_get_edges(ctx, nullptr, nullptr, nullptr, &m);
    if (rc) return rc;
    std::vector<int32_t> u((size_t)m), v((size_t)m);
    if (m > 0 && (rc = emst_get_edges(ctx, u.data(), v.data(), nullptr, &m))) return rc;
    double unit = 0;
    int64_t n = 0;
    int32_t dim = 0;
    if ((rc = emst_get_shape(ctx, &n, &dim))) return rc;
    std::vector<float> resc((size_t)(n * dim));
    std::vector<int32_t> lab((size_t)n);
    if (n > 0) {
        if ((rc = emst_get_rescaled(ctx, resc.data(), &unit))) return rc;
        if ((rc = emst_get_labels(ctx, lab.data()))) return rc;
    }

    // .links (NeuralNetEMST.h:277-300, BufferLink.h:266-279): every node,
    // its neighbours in ascending id order (the reference's order is the
    // order its atomics ran in)
    std::vector<std::vector<int32_t>> adj((size_t)n);
    for (int64_t e = 0; e < m; ++e) {
        adj[(size_t)u[(size_t)e]].push_back(v[(size_t)e]);
        adj[(size_t)v[(size_t)e]].push_back(u[(size_t)e]);
    }
    std::string o = grid_header(n);
    for (int64_t i = 0; i < n; ++i) {
        auto& a = adj[(size_t)i];
        std::sort(a.begin(), a.end());
        o += "\nNodeP = " + std::to_string(i) + " 0 NumLinks = " + std::to_string(a.size()) + " ";
        for (int32_t j : a) o += std::to_string(j) + " 0 ";
        o += "\n";
    }
    if (!write_file(base + ".links", o)) return EMST_E_IO;

    // .correspondenceMap: the host copy the reference writes is the one its
    // initialisation set, (-1, -1) per point (SolutionEMST.cu:464-466; the
    // rounds only update the device copy)
    o = grid_header(n);
    for (int64_t i = 0; i < n; ++i) o += "-1 -1  ";
    o += "\n\n";
    if (!write_file(base + ".correspondenceMap", o)) return EMST_E_IO;

    // .disjointSetMap: root per point as a byte offset (4 x root index)
    o = grid_header(n);
    for (int64_t i = 0; i < n; ++i) o += std::to_string(4ll * lab[(size_t)i]) + " ";
    o += "\n\n";
    if (!write_file(base + ".disjointSetMap", o)) return EMST_E_IO;

    // .points: the rescaled map as 3D points (z = 0 for a 2D solution)
    o = grid_header(n);
    for (int64_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            put_g(o, k < dim ? (double)resc[(size_t)(i * dim + k)] : 0.0);
            o += " ";
        }
        o += " ";
    }
    o += "\n\n";
    if (!write_file(base + ".points", o)) return EMST_E_IO;
    return EMST_OK;
}

}  // extern "C"
