// Host-side file formats around the device path:
//   * PNG -> RGBA8 with the transformations the reference requests from
//     libpng (main.cpp:29-109): 16-bit stripped to the high byte, palette
//     expanded, gray below 8 bits expanded, tRNS turned into alpha, 0xFF
//     filler as the 4th byte.  Decoded here with zlib; interlaced (Adam7)
//     images are de-interlaced as libpng's png_read_image does.
//   * the solidified lens OBJ (src/utils.cpp:176-260), heightmap.json
//     (main.cpp:111-135) and the SVG progress outputs (src/utils.cpp:262-344,
//     src/mesh.cpp:101-103,411-414,608-610), written with the same stream
//     formatting so that equal inputs give byte-equal files.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <new>
#include <vector>

#include "cd_api.h"

namespace {

uint32_t be32(const unsigned char* p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

struct PngInfo {
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = 0, interlace = 0;
    std::vector<unsigned char> plte;   // RGB triples
    std::vector<unsigned char> trns;   // palette alphas, or the gray/RGB key (16-bit BE)
    std::vector<unsigned char> idat;
};

int channels_of(int ctype) {
    switch (ctype) {
        case 0: return 1;   // gray
        case 2: return 3;   // RGB
        case 3: return 1;   // palette index
        case 4: return 2;   // gray + alpha
        case 6: return 4;   // RGBA
        default: return 0;
    }
}

bool valid_depth(int ctype, int d) {
    switch (ctype) {
        case 0: return d == 1 || d == 2 || d == 4 || d == 8 || d == 16;
        case 3: return d == 1 || d == 2 || d == 4 || d == 8;
        case 2: case 4: case 6: return d == 8 || d == 16;
        default: return false;
    }
}

int read_chunks(const std::vector<unsigned char>& f, PngInfo& pi) {
    static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) return CD_ERR_BAD_ARG;
    size_t p = 8;
    bool have_ihdr = false, have_iend = false;
    while (p + 12 <= f.size()) {
        const uint32_t len = be32(&f[p]);
        if (len > f.size() - p - 12) return CD_ERR_BAD_ARG;
        const unsigned char* type = &f[p + 4];
        const unsigned char* data = &f[p + 8];
        const uint32_t crc = be32(&f[p + 8 + len]);
        if (crc32(crc32(0L, Z_NULL, 0), type, 4 + len) != crc) return CD_ERR_BAD_ARG;
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len != 13) return CD_ERR_BAD_ARG;
            pi.w = be32(data);
            pi.h = be32(data + 4);
            pi.depth = data[8];
            pi.ctype = data[9];
            pi.interlace = data[12];
            if (data[10] != 0 || data[11] != 0 || pi.interlace > 1) return CD_ERR_BAD_ARG;
            if (!channels_of(pi.ctype) || !valid_depth(pi.ctype, pi.depth)) return CD_ERR_BAD_ARG;
            if (pi.w == 0 || pi.h == 0 || pi.w > (1u << 16) || pi.h > (1u << 16)) return CD_ERR_BAD_ARG;
            have_ihdr = true;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            if (len % 3 || len > 768) return CD_ERR_BAD_ARG;
            pi.plte.assign(data, data + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            pi.trns.assign(data, data + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            pi.idat.insert(pi.idat.end(), data, data + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            have_iend = true;
            break;
        } else if (!(type[0] & 0x20)) {
            return CD_ERR_BAD_ARG;  // unknown critical chunk
        }
        p += 12 + len;
    }
    if (!have_ihdr || !have_iend || pi.idat.empty()) return CD_ERR_BAD_ARG;
    if (pi.ctype == 3 && pi.plte.empty()) return CD_ERR_BAD_ARG;
    return CD_OK;
}

int paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    if (pb <= pc) return b;
    return c;
}

// undo the scanline filters of one (sub)image in place; rows of `rowbytes`
// bytes, each preceded by its filter byte
int unfilter(unsigned char* buf, uint32_t rows, size_t rowbytes, int bpp) {
    std::vector<unsigned char> prev(rowbytes, 0);
    for (uint32_t y = 0; y < rows; ++y) {
        unsigned char* row = buf + y * (rowbytes + 1);
        const int ft = row[0];
        unsigned char* r = row + 1;
        for (size_t i = 0; i < rowbytes; ++i) {
            const int a = i >= (size_t)bpp ? r[i - bpp] : 0;
            const int b = prev[i];
            const int c = i >= (size_t)bpp ? prev[i - bpp] : 0;
            int v = r[i];
            switch (ft) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) >> 1; break;
                case 4: v += paeth(a, b, c); break;
                default: return CD_ERR_BAD_ARG;
            }
            r[i] = (unsigned char)v;
        }
        std::memcpy(prev.data(), r, rowbytes);
    }
    return CD_OK;
}

// sample s of a row at bit depth d (< 8: packed MSB first; 16: high byte)
inline int sample(const unsigned char* r, size_t s, int d) {
    if (d == 8) return r[s];
    if (d == 16) return r[2 * s];
    const int per = 8 / d;
    const int byte = r[s / per];
    const int shift = 8 - d * (int)(s % per + 1);
    return (byte >> shift) & ((1 << d) - 1);
}
inline int sample16(const unsigned char* r, size_t s) { return (r[2 * s] << 8) | r[2 * s + 1]; }

// one decoded pixel -> RGBA8 with the reference's libpng transformations
void to_rgba(const PngInfo& pi, const unsigned char* r, uint32_t x, unsigned char* o) {
    const int d = pi.depth;
    const bool key = !pi.trns.empty();
    switch (pi.ctype) {
        case 0: {
            int g = sample(r, x, d);
            bool transparent = false;
            if (key && pi.trns.size() >= 2) {
                const int k = (pi.trns[0] << 8) | pi.trns[1];
                transparent = (d == 16 ? sample16(r, x) : g) == k;
            }
            if (d < 8) g = g * (255 / ((1 << d) - 1));
            o[0] = o[1] = o[2] = (unsigned char)g;
            o[3] = transparent ? 0 : 255;
            break;
        }
        case 2: {
            bool transparent = false;
            if (key && pi.trns.size() >= 6) {
                const int kr = (pi.trns[0] << 8) | pi.trns[1], kg = (pi.trns[2] << 8) | pi.trns[3],
                          kb = (pi.trns[4] << 8) | pi.trns[5];
                if (d == 16)
                    transparent = sample16(r, 3 * x) == kr && sample16(r, 3 * x + 1) == kg &&
                                  sample16(r, 3 * x + 2) == kb;
                else
                    transparent = sample(r, 3 * x, d) == kr && sample(r, 3 * x + 1, d) == kg &&
                                  sample(r, 3 * x + 2, d) == kb;
            }
            for (int c = 0; c < 3; ++c) o[c] = (unsigned char)sample(r, 3 * x + c, d);
            o[3] = transparent ? 0 : 255;
            break;
        }
        case 3: {
            const int i = sample(r, x, d);
            const size_t n = pi.plte.size() / 3;
            for (int c = 0; c < 3; ++c) o[c] = (size_t)i < n ? pi.plte[3 * i + c] : 0;
            o[3] = (size_t)i < pi.trns.size() ? pi.trns[i] : 255;
            break;
        }
        case 4: {
            const int g = sample(r, 2 * x, d);
            o[0] = o[1] = o[2] = (unsigned char)g;
            o[3] = (unsigned char)sample(r, 2 * x + 1, d);
            break;
        }
        default: {  // 6
            for (int c = 0; c < 4; ++c) o[c] = (unsigned char)sample(r, 4 * x + c, d);
        }
    }
}

int decode_png(const char* path, std::vector<unsigned char>& rgba, uint32_t& w, uint32_t& h) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return CD_ERR_BAD_ARG;
    std::vector<unsigned char> f((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    PngInfo pi;
    int st = read_chunks(f, pi);
    if (st != CD_OK) return st;
    w = pi.w;
    h = pi.h;
    const int ch = channels_of(pi.ctype);
    const int bits = ch * pi.depth;
    const int bpp = std::max(1, bits / 8);
    // Adam7 passes (start x, start y, step x, step y); one pass when not interlaced
    static const int A7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                 {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    const int npass = pi.interlace ? 7 : 1;
    size_t total = 0;
    for (int p = 0; p < npass; ++p) {
        const uint32_t sx = pi.interlace ? A7[p][0] : 0, sy = pi.interlace ? A7[p][1] : 0;
        const uint32_t dx = pi.interlace ? A7[p][2] : 1, dy = pi.interlace ? A7[p][3] : 1;
        const uint32_t pw = w > sx ? (w - sx + dx - 1) / dx : 0;
        const uint32_t ph = h > sy ? (h - sy + dy - 1) / dy : 0;
        if (pw && ph) total += (size_t)ph * (((size_t)pw * bits + 7) / 8 + 1);
    }
    // Refuse sizes before allocating: deflate expands at most ~1032:1, so a
    // header claiming more filtered bytes than the IDAT data can inflate to is
    // corrupt (or hostile); and the decoded image is capped at 2^28 pixels
    // (1 GiB of RGBA), far above any target this path is used for.
    if ((uint64_t)w * h > (1ull << 28) || total > (uint64_t)pi.idat.size() * 1032 + 1024 ||
        total > 0xFFFFFFFFull)
        return CD_ERR_BAD_ARG;
    std::vector<unsigned char> raw(total);
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) return CD_ERR_BAD_ARG;
    zs.next_in = pi.idat.data();
    zs.avail_in = (uInt)pi.idat.size();
    zs.next_out = raw.data();
    zs.avail_out = (uInt)raw.size();
    const int zr = inflate(&zs, Z_FINISH);
    const size_t got = raw.size() - zs.avail_out;
    inflateEnd(&zs);
    if ((zr != Z_STREAM_END && zr != Z_BUF_ERROR) || got != total) return CD_ERR_BAD_ARG;
    rgba.assign((size_t)w * h * 4, 0);
    size_t off = 0;
    for (int p = 0; p < npass; ++p) {
        const uint32_t sx = pi.interlace ? A7[p][0] : 0, sy = pi.interlace ? A7[p][1] : 0;
        const uint32_t dx = pi.interlace ? A7[p][2] : 1, dy = pi.interlace ? A7[p][3] : 1;
        const uint32_t pw = w > sx ? (w - sx + dx - 1) / dx : 0;
        const uint32_t ph = h > sy ? (h - sy + dy - 1) / dy : 0;
        if (!pw || !ph) continue;
        const size_t rowbytes = ((size_t)pw * bits + 7) / 8;
        st = unfilter(raw.data() + off, ph, rowbytes, bpp);
        if (st != CD_OK) return st;
        for (uint32_t y = 0; y < ph; ++y) {
            const unsigned char* r = raw.data() + off + y * (rowbytes + 1) + 1;
            for (uint32_t x = 0; x < pw; ++x)
                to_rgba(pi, r, x, &rgba[(((size_t)(sy + y * dy)) * w + (sx + x * dx)) * 4]);
        }
        off += (size_t)ph * (rowbytes + 1);
    }
    return CD_OK;
}

void find_perimeter(int nx, int ny, std::vector<int>& out) {
    // the reference's walk (utils.cpp:176-196): top row, right column, bottom
    // row right to left, left column bottom to top (corners repeat as there)
    for (int i = 0; i < nx; ++i) out.push_back(i);
    for (int i = nx - 1; i < nx * ny; i += nx) out.push_back(i);
    for (int i = nx * (ny - 1) + nx - 1; i > nx * (ny - 1) - 1; --i) out.push_back(i);
    for (int i = nx * (ny - 1) - nx; i > nx - 1; i -= nx) out.push_back(i);
}

struct Tri {
    int a, b, c;
};

std::vector<Tri> triangles(int nx, int ny) {
    // structured triangulation (mesh.cpp:57-63)
    std::vector<Tri> t;
    t.reserve(2 * (size_t)(nx - 1) * (ny - 1));
    for (int i = 0; i < ny - 1; ++i)
        for (int j = 0; j < nx - 1; ++j) {
            const int idx = i * nx + j;
            t.push_back({idx, idx + 1, idx + nx});
            t.push_back({idx + nx, idx + 1, idx + nx + 1});
        }
    return t;
}

int get_field(cd_design* d, int field, std::vector<double>& out) {
    out.resize(cd_design_field_size(d, field));
    return cd_design_get(d, field, out.data(), out.size());
}

// "x,y" of one point in the reference's SVG units, same operation order
// (utils.cpp:280-281): (x / width) * 1000, ((y / height) * 1000) * (height / width)
std::string svg_xy(const double* p, double width, double height) {
    return std::to_string((p[0] / width) * 1000.0f) + "," +
           std::to_string((p[1] / height) * 1000.0f * (height / width));
}

}  // namespace

extern "C" {

int cd_png_read_rgba8(const char* path, int* width, int* height, unsigned char* rgba,
                      size_t capacity) {
    if (!path || !width || !height) return CD_ERR_BAD_ARG;
    std::vector<unsigned char> img;
    uint32_t w = 0, h = 0;
    int st = CD_OK;
    try {  // no C++ exception may leave the C ABI
        st = decode_png(path, img, w, h);
    } catch (const std::bad_alloc&) {
        return CD_ERR_OOM;
    } catch (...) {
        return CD_ERR_BAD_ARG;
    }
    if (st != CD_OK) return st;
    *width = (int)w;
    *height = (int)h;
    if (!rgba) return CD_OK;
    if (capacity < img.size()) return CD_ERR_BAD_ARG;
    std::memcpy(rgba, img.data(), img.size());
    return CD_OK;
}

int cd_obj_write_solid(const char* path, const double* front, const double* back, int res_x,
                       int res_y, double thickness, double width, double height) {
    if (!path || !front || !back || res_x < 2 || res_y < 2) return CD_ERR_BAD_ARG;
    std::ofstream file(path);
    if (!file.is_open()) return CD_ERR_BAD_ARG;
    const int n = res_x * res_y;
    double min_h = 0;
    for (int i = 0; i < n; ++i) min_h = std::min(min_h, front[3 * i + 2]);
    file << "# Generated by the software algorithm written by Dylan Missuwe" << "\n";
    file << "# The algorithm used to create the lens is based on the paper "
         << "\"Poisson-Based Continuous Surface Generation for Goal-Based Caustics\" "
         << "by Yue et al (2014)" << "\n";
    for (int i = 0; i < n; ++i)
        file << "v " << width - front[3 * i] << " " << height - front[3 * i + 1] << " "
             << -(front[3 * i + 2]) << "\n";
    for (int i = 0; i < n; ++i)
        file << "v " << width - back[3 * i] << " " << height - back[3 * i + 1] << " "
             << -thickness - min_h << "\n";
    const std::vector<Tri> tris = triangles(res_x, res_y);
    for (const Tri& t : tris) file << "f " << t.a + 1 << " " << t.b + 1 << " " << t.c + 1 << "\n";
    for (const Tri& t : tris)
        file << "f " << t.a + n + 1 << " " << t.c + n + 1 << " " << t.b + n + 1 << "\n";
    std::vector<int> per;
    find_perimeter(res_x, res_y, per);
    for (size_t i = 0; i < per.size(); ++i) {
        const int top = per[i], bottom = per[i] + n;
        const int ntop = per[(i + 1) % per.size()], nbottom = ntop + n;
        file << "f " << top + 1 << " " << bottom + 1 << " " << nbottom + 1 << "\n";
        file << "f " << top + 1 << " " << nbottom + 1 << " " << ntop + 1 << "\n";
    }
    return file.good() ? CD_OK : CD_ERR_BAD_ARG;
}

int cd_json_write_heightmap(const char* path, const double* h, int res_x, int res_y) {
    if (!path || !h || res_x < 1 || res_y < 1) return CD_ERR_BAD_ARG;
    std::ofstream out(path);
    if (!out.is_open()) return CD_ERR_BAD_ARG;
    out << "[\n";
    for (int i = 0; i < res_y; ++i) {
        out << "  [";
        for (int j = 0; j < res_x; ++j) {
            out << h[(size_t)i * res_x + j];
            if (j < res_x - 1) out << ", ";
        }
        out << "]";
        out << (i < res_y - 1 ? ",\n" : "\n");
    }
    out << "]\n";
    return out.good() ? CD_OK : CD_ERR_BAD_ARG;
}

int cd_svg_write_grid(const char* path, const double* points, int stride, int res_x, int res_y,
                      double width, double height, double stroke_width) {
    if (!path || !points || stride < 2 || res_x < 1 || res_y < 1) return CD_ERR_BAD_ARG;
    std::ofstream svg(path, std::ios::out);
    if (!svg.is_open()) return CD_ERR_BAD_ARG;
    svg << "<?xml version=\"1.0\" encoding=\"UTF-8\" ?>\n";
    svg << "<svg width=\"1000\" height=\"" << 1000.0f * (height / width)
        << "\" xmlns=\"http://www.w3.org/2000/svg\" xmlns:xlink=\"http://www.w3.org/1999/xlink\">\n";
    svg << "<rect width=\"100%\" height=\"100%\" fill=\"white\"/>\n";
    auto pt = [&](int idx) { return svg_xy(points + (size_t)idx * stride, width, height); };
    for (int j = 0; j < res_y; ++j) {
        std::string d = "M";
        for (int i = 0; i < res_x; ++i) {
            d += pt(i + j * res_x);
            if (i < res_x - 1) d += "L";
        }
        svg << "<path d=\"" << d << "\" fill=\"none\" stroke=\"black\" stroke-width=\""
            << stroke_width << "\"/>\n";
    }
    for (int j = 0; j < res_x; ++j) {
        std::string d = "M";
        for (int i = 0; i < res_y; ++i) {
            d += pt(j + i * res_x);
            // the reference compares against res_x here (utils.cpp:292)
            if (i < res_x - 1) d += "L";
        }
        svg << "<path d=\"" << d << "\" fill=\"none\" stroke=\"black\" stroke-width=\""
            << stroke_width << "\"/>\n";
    }
    svg << "</svg>\n";
    return svg.good() ? CD_OK : CD_ERR_BAD_ARG;
}

int cd_svg_write_triangles(const char* path, const double* points, int stride, int res_x,
                           int res_y, double width, double height, double stroke_width) {
    if (!path || !points || stride < 2 || res_x < 2 || res_y < 2) return CD_ERR_BAD_ARG;
    std::ofstream svg(path, std::ios::out);
    if (!svg.is_open()) return CD_ERR_BAD_ARG;
    svg << "<?xml version=\"1.0\" encoding=\"UTF-8\" ?>\n";
    svg << "<svg width=\"1000\" height=\"" << 1000.0f * (height / width)
        << "\" xmlns=\"http://www.w3.org/2000/svg\" xmlns:xlink=\"http://www.w3.org/1999/xlink\">\n";
    svg << "<rect width=\"100%\" height=\"100%\" fill=\"white\"/>\n";
    for (const Tri& t : triangles(res_x, res_y)) {
        const int v[3] = {t.a, t.b, t.c};
        std::string d = "M";
        for (int j = 0; j < 3; ++j) {
            d += svg_xy(points + (size_t)v[j] * stride, width, height);
            if (j < 2) d += "L";
        }
        d += "Z";
        svg << "<path d=\"" << d << "\" fill=\"none\" stroke=\"black\" stroke-width=\""
            << stroke_width << "\"/>\n";
    }
    svg << "</svg>\n";
    return svg.good() ? CD_OK : CD_ERR_BAD_ARG;
}

int cd_design_save_solid_obj(cd_design* d, const char* path, int front) {
    if (!d || !path || (front != CD_OBJ_SOURCE && front != CD_OBJ_TARGET)) return CD_ERR_BAD_ARG;
    cd_design_params p;
    int st = cd_design_get_params(d, &p);
    if (st != CD_OK) return st;
    std::vector<double> src, tgt;
    if ((st = get_field(d, CD_FIELD_SOURCE_POINTS, src)) != CD_OK) return st;
    if (front == CD_OBJ_TARGET && (st = get_field(d, CD_FIELD_TARGET_POINTS, tgt)) != CD_OK)
        return st;
    return cd_obj_write_solid(path, front == CD_OBJ_TARGET ? tgt.data() : src.data(), src.data(),
                              p.mesh_res_x, p.mesh_res_y, p.thickness, p.width, p.height);
}

int cd_design_save_heightmap_json(cd_design* d, const char* path) {
    if (!d || !path) return CD_ERR_BAD_ARG;
    cd_design_params p;
    int st = cd_design_get_params(d, &p);
    if (st != CD_OK) return st;
    std::vector<double> h;
    if ((st = get_field(d, CD_FIELD_H, h)) != CD_OK) return st;
    return cd_json_write_heightmap(path, h.data(), p.res_x, p.res_y);
}

int cd_design_export_svg(cd_design* d, const char* path, int what, double stroke_width) {
    if (!d || !path) return CD_ERR_BAD_ARG;
    cd_design_params p;
    int st = cd_design_get_params(d, &p);
    if (st != CD_OK) return st;
    std::vector<double> pts;
    switch (what) {
        case CD_SVG_PARAMETERIZATION:
            if ((st = get_field(d, CD_FIELD_TARGET_POINTS, pts)) != CD_OK) return st;
            return cd_svg_write_grid(path, pts.data(), 3, p.mesh_res_x, p.mesh_res_y, p.width,
                                     p.height, stroke_width);
        case CD_SVG_INVERTED_MAP:
            if ((st = get_field(d, CD_FIELD_INVERSE_MAP, pts)) != CD_OK) return st;
            return cd_svg_write_grid(path, pts.data(), 2, p.mesh_res_x, p.mesh_res_y, p.width,
                                     p.height, stroke_width);
        case CD_SVG_TRIANGLES:
            if ((st = get_field(d, CD_FIELD_TARGET_POINTS, pts)) != CD_OK) return st;
            return cd_svg_write_triangles(path, pts.data(), 3, p.mesh_res_x, p.mesh_res_y,
                                          p.width, p.height, stroke_width);
        default:
            return CD_ERR_BAD_ARG;
    }
}

}  // extern "C"
