// lb_deck.cpp -- the reference's input decks and VTK output, host side.
//
//  * parameter files (.dat): thirteen lines, one value each, in the order of
//    Input::loadParameters (input.cpp:51-97); lines past the end read as 0
//    (the 3D sample decks stop after updateSteps);
//  * 2D solids from a BMP: black pixels (0,0,0) are solid, row 0 is the top
//    row of the image (Input::loadSolids, input.cpp:99-116); uncompressed
//    BMPs with 1, 4, 8, 24 or 32 bits per pixel, bottom-up or top-down;
//  * 3D solids from an STL (binary or ASCII): the mesh is scaled so its
//    longest extent spans max_dim voxels and each voxel whose centre lies
//    inside the closed surface is solid (the reference calls cvmlcpp's
//    voxelizer for this, input.cpp:118-142; this is a scan-line parity
//    voxelizer of our own);
//  * VTK image data: the ASCII .vti piece per subdomain and the .pvti index
//    of Output::saveVTK / saveParVTK (output.cpp:52-155), same file names,
//    extents and value layout, so existing ParaView state files keep working.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/lattibol_amd.h"

namespace {

// One line of a parameter file as a number; empty / missing / malformed -> 0.
double parse_number(const std::string &line, bool integer) {
    const char *s = line.c_str();
    while (*s == ' ' || *s == '\t') s++;
    if (!*s) return 0.0;
    char *end = nullptr;
    double v = integer ? (double)std::strtol(s, &end, 10) : std::strtod(s, &end);
    if (end == s) return 0.0;
    while (*end == ' ' || *end == '\t' || *end == '\r' || *end == '\n') end++;
    if (*end) return 0.0;   // trailing garbage: the whole line fails to convert
    return v;
}

uint32_t rd32(const unsigned char *p) { return p[0] | (p[1] << 8) | (p[2] << 16) | ((uint32_t)p[3] << 24); }
uint16_t rd16(const unsigned char *p) { return (uint16_t)(p[0] | (p[1] << 8)); }

bool read_file(const char *path, std::vector<unsigned char> &buf) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    buf.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
}

struct Tri { float v[3][3]; };

bool read_stl(const char *path, std::vector<Tri> &tris) {
    std::vector<unsigned char> buf;
    if (!read_file(path, buf)) return false;
    tris.clear();
    // binary: 80-byte header, uint32 count, 50 bytes per facet
    if (buf.size() >= 84) {
        const uint32_t n = rd32(&buf[80]);
        if ((uint64_t)84 + (uint64_t)n * 50 == buf.size()) {
            tris.resize(n);
            for (uint32_t t = 0; t < n; t++) {
                const unsigned char *p = &buf[84 + (size_t)t * 50 + 12];
                for (int k = 0; k < 3; k++)
                    for (int a = 0; a < 3; a++) std::memcpy(&tris[t].v[k][a], p + 4 * (3 * k + a), 4);
            }
            return true;
        }
    }
    // ASCII: "vertex x y z" lines, three per facet
    std::istringstream in(std::string(buf.begin(), buf.end()));
    std::string word;
    Tri cur{};
    int k = 0;
    while (in >> word) {
        if (word != "vertex") continue;
        float x, y, z;
        if (!(in >> x >> y >> z)) return false;
        cur.v[k][0] = x; cur.v[k][1] = y; cur.v[k][2] = z;
        if (++k == 3) { tris.push_back(cur); k = 0; }
    }
    return !tris.empty();
}

}  // namespace

extern "C" int lb_deck_read_params(const char *dat_path, lb_deck_params *out) {
    if (!dat_path || !out) return LB_ERR_INVALID;
    std::ifstream f(dat_path);
    if (!f) return LB_ERR_INVALID;
    std::string lines[13];
    for (auto &l : lines)
        if (!std::getline(f, l)) l.clear();
    lb_deck_params p{};
    p.device_count = (int)parse_number(lines[0], true);
    const int md = (int)parse_number(lines[1], true);
    p.max_dim = md > 1 ? md : 1;   // <= 1: a 2D deck with a BMP mask
    p.ux = (float)parse_number(lines[2], false);
    p.uy = (float)parse_number(lines[3], false);
    p.uz = (float)parse_number(lines[4], false);
    p.rho = (float)parse_number(lines[5], false);
    p.viscosity = (float)parse_number(lines[6], false);
    p.tau_inv = (float)(1. / (3. * (double)p.viscosity + 0.5));   // input.cpp:168-171
    p.update_steps = (int)parse_number(lines[7], true);
    p.fx = (float)parse_number(lines[8], false);
    p.fy = (float)parse_number(lines[9], false);
    p.fz = (float)parse_number(lines[10], false);
    p.G = (float)parse_number(lines[11], false);
    p.Gads = (float)parse_number(lines[12], false);
    p.min_u = 0.0f;
    p.max_u = p.ux * 1.5f;
    p.min_rho = p.rho * 0.98f;
    p.max_rho = p.rho * 1.02f;
    *out = p;
    return 0;
}

extern "C" int lb_deck_read_bmp(const char *bmp_path, int *width, int *height, uint8_t *solids) {
    if (!bmp_path || !width || !height) return LB_ERR_INVALID;
    std::vector<unsigned char> b;
    if (!read_file(bmp_path, b) || b.size() < 54 || b[0] != 'B' || b[1] != 'M') return LB_ERR_INVALID;
    const uint32_t off = rd32(&b[10]), hdr = rd32(&b[14]);
    if (hdr < 40) return LB_ERR_UNSUPPORTED;   // OS/2 core headers
    const int32_t w = (int32_t)rd32(&b[18]);
    const int32_t h_signed = (int32_t)rd32(&b[22]);
    const int bpp = rd16(&b[28]);
    const uint32_t comp = rd32(&b[30]);
    uint32_t ncol = rd32(&b[46]);
    if (w <= 0 || h_signed == 0 || comp != 0) return LB_ERR_UNSUPPORTED;   // BI_RGB only
    // a lattice edge beyond 65,536 nodes is no deck; also keeps -h_signed
    // defined and stride * h far from overflow
    if (w > (1 << 16) || h_signed > (1 << 16) || h_signed < -(1 << 16)) return LB_ERR_UNSUPPORTED;
    if (bpp != 1 && bpp != 4 && bpp != 8 && bpp != 24 && bpp != 32) return LB_ERR_UNSUPPORTED;
    const bool top_down = h_signed < 0;
    const int h = top_down ? -h_signed : h_signed;
    if (bpp <= 8 && ncol == 0) ncol = 1u << bpp;
    const size_t pal = 14 + (size_t)hdr;
    const size_t stride = (((size_t)w * bpp + 31) / 32) * 4;
    if (bpp <= 8 && pal + 4 * (size_t)ncol > b.size()) return LB_ERR_INVALID;
    if ((size_t)off > b.size() || stride * (size_t)h > b.size() - (size_t)off) return LB_ERR_INVALID;
    *width = w;
    *height = h;
    if (!solids) return 0;
    for (int y = 0; y < h; y++) {
        const unsigned char *row = &b[off + stride * (size_t)(top_down ? y : h - 1 - y)];
        for (int x = 0; x < w; x++) {
            unsigned r, g, bl;
            if (bpp <= 8) {
                const unsigned bit = (unsigned)x * bpp;
                const unsigned idx = (row[bit / 8] >> (8 - bpp - bit % 8)) & ((1u << bpp) - 1);
                if (idx >= ncol) return LB_ERR_INVALID;
                const unsigned char *c = &b[pal + 4 * (size_t)idx];   // B, G, R, reserved
                bl = c[0]; g = c[1]; r = c[2];
            } else {
                const unsigned char *c = row + (size_t)x * (bpp / 8);
                bl = c[0]; g = c[1]; r = c[2];
            }
            solids[x + (size_t)w * y] = (r == 0 && g == 0 && bl == 0) ? 1 : 0;
        }
    }
    return 0;
}

extern "C" int lb_deck_voxelize_stl(const char *stl_path, int max_dim, int *width, int *height, int *depth,
                                    uint8_t *solids) {
    if (!stl_path || !width || !height || !depth || max_dim < 1) return LB_ERR_INVALID;
    std::vector<Tri> tris;
    if (!read_stl(stl_path, tris) || tris.empty()) return LB_ERR_INVALID;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (const Tri &t : tris)
        for (int k = 0; k < 3; k++)
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], (double)t.v[k][a]);
                hi[a] = std::max(hi[a], (double)t.v[k][a]);
            }
    const double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
    if (!(ext > 0)) return LB_ERR_INVALID;
    const double vs = ext / max_dim;   // voxel edge in mesh units: the longest extent spans max_dim voxels
    int n[3];
    for (int a = 0; a < 3; a++) n[a] = std::max(1, (int)std::ceil((hi[a] - lo[a]) / vs - 1e-6));
    *width = n[0];
    *height = n[1];
    *depth = n[2];
    if (!solids) return 0;
    std::fill(solids, solids + (size_t)n[0] * n[1] * n[2], 0);
    // scan lines along z through the voxel centres (x, y); a tiny offset keeps
    // the line off mesh edges that lie exactly on a centre line
    std::vector<double> zs;
    for (int y = 0; y < n[1]; y++)
        for (int x = 0; x < n[0]; x++) {
            const double px = lo[0] + (x + 0.5) * vs + 1.3e-7 * vs, py = lo[1] + (y + 0.5) * vs + 0.7e-7 * vs;
            zs.clear();
            for (const Tri &t : tris) {
                const double ax = t.v[0][0], ay = t.v[0][1], bx = t.v[1][0], by = t.v[1][1], cx = t.v[2][0],
                             cy = t.v[2][1];
                const double d = (by - cy) * (ax - cx) + (cx - bx) * (ay - cy);
                if (d == 0) continue;   // triangle seen edge-on from the scan direction
                const double l1 = ((by - cy) * (px - cx) + (cx - bx) * (py - cy)) / d;
                const double l2 = ((cy - ay) * (px - cx) + (ax - cx) * (py - cy)) / d;
                const double l3 = 1 - l1 - l2;
                if (l1 < 0 || l2 < 0 || l3 < 0) continue;
                zs.push_back(l1 * t.v[0][2] + l2 * t.v[1][2] + l3 * t.v[2][2]);
            }
            if (zs.size() < 2) continue;
            std::sort(zs.begin(), zs.end());
            for (size_t k = 0; k + 1 < zs.size(); k += 2)
                for (int z = 0; z < n[2]; z++) {
                    const double pz = lo[2] + (z + 0.5) * vs;
                    if (pz > zs[k] && pz < zs[k + 1]) solids[x + (size_t)n[0] * (y + (size_t)n[1] * z)] = 1;
                }
        }
    return 0;
}

// File names r<id4>.pvti and r<id4>p<dev2>.vti as in output.cpp:52-80 / 95-115.
static std::string vtk_name(const char *dir, int vtk_id, int dev, bool piece) {
    char name[32];
    if (piece) std::snprintf(name, sizeof name, "r%04dp%02d.vti", vtk_id % 10000, dev % 100);
    else std::snprintf(name, sizeof name, "r%04d.pvti", vtk_id % 10000);
    std::string d = dir && *dir ? dir : ".";
    return d + "/" + name;
}

extern "C" int lb_write_pvti(const char *dir, int vtk_id, int device_count, int width, int height, int length) {
    if (device_count < 1 || width < 1 || height < 1 || length < 1 || vtk_id < 0) return LB_ERR_INVALID;
    std::ofstream f(vtk_name(dir, vtk_id, 0, false));
    if (!f) return LB_ERR_INVALID;
    f << "<VTKFile type=\"PImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
      << "  <PImageData WholeExtent=\"0 " << width * device_count - 1 << " 0 " << height - 1 << " 0 " << length - 1
      << "\" GhostLevel=\"#\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n"
      << "    <PPointData Scalars=\"Pressure\" Vectors=\"Velocity\">\n"
      << "      <DataArray type=\"Float32\" Name=\"Pressure\" format=\"ascii\"/>\n"
      << "      <DataArray type=\"Float32\" Name=\"Velocity\" NumberOfComponents=\"3\" format=\"ascii\"/>\n"
      << "    </PPointData>\n    <PCellData>\n    </PCellData>\n";
    for (int i = 0; i < device_count; i++) {
        std::string src = vtk_name(nullptr, vtk_id, i, true).substr(2);   // bare file name
        f << "    <Piece Extent=\"" << i * width << " " << (i + 1) * width << " 0 " << height - 1 << " 0 " << length - 1
          << "\" Source=\"" << src << "\"/>\n";
    }
    f << "  </PImageData>\n</VTKFile>\n";
    return f ? 0 : LB_ERR_INVALID;
}

// One subdomain's piece: rho ("Pressure") and (ux, uy, uz) ("Velocity") with
// width x height x length nodes, x fastest.  The piece extent spans width + 1
// points in x (adjacent pieces share a face), so each row ends with one pad
// value; a 2D lattice (length 1) writes (ux, -uy, 0), physical y pointing up
// while rows grow downward.  Default ostream formatting (6 significant
// digits), as the reference writes.
extern "C" int lb_write_vti(const char *dir, int vtk_id, int device_id, int width, int height, int length,
                            const float *ux, const float *uy, const float *uz, const float *rho) {
    if (!ux || !uy || !rho || (length > 1 && !uz) || width < 1 || height < 1 || length < 1 || device_id < 0 ||
        vtk_id < 0)
        return LB_ERR_INVALID;
    std::ofstream f(vtk_name(dir, vtk_id, device_id, true));
    if (!f) return LB_ERR_INVALID;
    const size_t n = (size_t)width * height * length;
    f << "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
      << "  <ImageData WholeExtent=\"" << device_id * width << " " << (device_id + 1) * width << " 0 " << height - 1
      << " 0 " << length - 1 << "\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n"
      << "  <Piece Extent=\"" << device_id * width << " " << (device_id + 1) * width << " 0 " << height - 1 << " 0 "
      << length - 1 << "\">\n"
      << "    <PointData Scalars=\"Pressure\" Vectors=\"Velocity\">\n"
      << "      <DataArray type=\"Float32\" Name=\"Pressure\" format=\"ascii\" >\n      ";
    for (size_t i = 0; i < n; i++) {
        f << rho[i] << " ";
        if ((i + 1) % width == 0) f << "0\n      ";
    }
    f << "</DataArray>\n"
      << "      <DataArray type=\"Float32\" Name=\"Velocity\" NumberOfComponents=\"3\" format=\"ascii\">\n"
      << "      ";
    for (size_t i = 0; i < n; i++) {
        f << ux[i] << " ";
        if (length > 1) f << uy[i] << " " << uz[i] << "   ";
        else f << -uy[i] << " 0   ";
        if ((i + 1) % width == 0) f << "0 0 0\n      ";
    }
    f << "</DataArray>\n    </PointData>\n    <CellData>\n    </CellData>\n  </Piece>\n  </ImageData>\n</VTKFile>\n";
    return f ? 0 : LB_ERR_INVALID;
}
