// Graph.h -- HeatPlot (Graph.h:7-99 of the reference) without matplotlib.
// stratiflow.h includes it after the field classes, where the reference's
// IMEXRK.h includes its Graph.h.
//
// The field is resampled exactly as the reference's HeatPlot does it: two
// periods in x at 50 pixels per unit length, z in [-6, 6) at 50 pixels per
// unit length, bilinear in x (periodic) and along the stretched vertical grid
// of the field's boundary condition, first row at the top.  The reference then
// draws the array with matplotlib's imshow (axes, default colormap); here it is
// written as an 8-bit RGB PNG of the array itself, the data range mapped onto a
// dark-blue -> white -> dark-red ramp, deflated by the system's libz when it
// can be loaded (else stored uncompressed).  HeatPlotData returns the
// resampled array for callers that draw it themselves.
#pragma once

#include <dlfcn.h>

#include <algorithm>
#include <cstdint>
#include <fstream>
#include <string>
#include <vector>

namespace sfacade {

// zlib's compress2, looked up at run time so that the facade needs no link flag
typedef int (*compress2_fn)(unsigned char*, unsigned long*, const unsigned char*, unsigned long, int);
typedef unsigned long (*compress_bound_fn)(unsigned long);
inline bool zlib_compress(const std::vector<unsigned char>& in, std::vector<unsigned char>& out) {
    static void* lib = dlopen("libz.so.1", RTLD_LAZY | RTLD_LOCAL);
    if (!lib) return false;
    static compress2_fn c2 = reinterpret_cast<compress2_fn>(dlsym(lib, "compress2"));
    static compress_bound_fn cb = reinterpret_cast<compress_bound_fn>(dlsym(lib, "compressBound"));
    if (!c2 || !cb) return false;
    unsigned long n = cb(in.size());
    out.resize(n);
    if (c2(out.data(), &n, in.data(), in.size(), 6) != 0) return false;
    out.resize(n);
    return true;
}

inline uint32_t png_crc(const unsigned char* p, size_t n, uint32_t c = 0xffffffffu) {
    for (size_t i = 0; i < n; ++i) {
        c ^= p[i];
        for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (0xedb88320u & (0u - (c & 1u)));
    }
    return c;
}

inline void png_u32(std::vector<unsigned char>& v, uint32_t x) {
    v.push_back(static_cast<unsigned char>(x >> 24));
    v.push_back(static_cast<unsigned char>(x >> 16));
    v.push_back(static_cast<unsigned char>(x >> 8));
    v.push_back(static_cast<unsigned char>(x));
}

inline void png_chunk(std::ofstream& out, const char* type, const std::vector<unsigned char>& data) {
    std::vector<unsigned char> c;
    png_u32(c, uint32_t(data.size()));
    c.insert(c.end(), type, type + 4);
    c.insert(c.end(), data.begin(), data.end());
    const uint32_t crc = png_crc(c.data() + 4, c.size() - 4) ^ 0xffffffffu;
    png_u32(c, crc);
    out.write(reinterpret_cast<const char*>(c.data()), std::streamsize(c.size()));
}

// an RGB image (rows top to bottom, 3 bytes per pixel) as a PNG file
inline void write_png(const std::string& filename, int width, int height, const std::vector<unsigned char>& rgb) {
    std::vector<unsigned char> raw;  // each row: filter type 0, then the pixels
    raw.reserve(size_t(height) * (3 * width + 1));
    for (int r = 0; r < height; ++r) {
        raw.push_back(0);
        raw.insert(raw.end(), rgb.begin() + size_t(r) * 3 * width, rgb.begin() + size_t(r + 1) * 3 * width);
    }
    std::vector<unsigned char> z;
    if (!zlib_compress(raw, z)) {
        // zlib stream of stored deflate blocks (RFC 1950/1951), Adler-32 at the end
        z = {0x78, 0x01};
        uint32_t a = 1, b = 0;
        for (unsigned char x : raw) {
            a = (a + x) % 65521u;
            b = (b + a) % 65521u;
        }
        for (size_t pos = 0;;) {
            const size_t n = std::min<size_t>(65535, raw.size() - pos);
            const bool last = pos + n == raw.size();
            z.push_back(last ? 1 : 0);
            z.push_back(static_cast<unsigned char>(n & 0xff));
            z.push_back(static_cast<unsigned char>(n >> 8));
            z.push_back(static_cast<unsigned char>(~n & 0xff));
            z.push_back(static_cast<unsigned char>((~n >> 8) & 0xff));
            z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + n);
            pos += n;
            if (last) break;
        }
        png_u32(z, (b << 16) | a);
    }
    std::ofstream out(filename, std::ios::binary);
    const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    out.write(reinterpret_cast<const char*>(sig), 8);
    std::vector<unsigned char> ihdr;
    png_u32(ihdr, uint32_t(width));
    png_u32(ihdr, uint32_t(height));
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8 bits, RGB, deflate, filter 0, no interlace
    png_chunk(out, "IHDR", ihdr);
    png_chunk(out, "IDAT", z);
    png_chunk(out, "IEND", {});
}

// t in [0, 1] -> dark blue (0) -> white (0.5) -> dark red (1)
inline void ramp(double t, unsigned char* rgb) {
    t = std::min(1.0, std::max(0.0, t));
    double r, g, b;
    if (t < 0.5) {
        const double s = t / 0.5;
        r = 0.05 + 0.95 * s;
        g = 0.15 + 0.85 * s;
        b = 0.45 + 0.55 * s;
    } else {
        const double s = (t - 0.5) / 0.5;
        r = 1.0 - 0.45 * s;
        g = 1.0 - 0.95 * s;
        b = 1.0 - 0.95 * s;
    }
    rgb[0] = static_cast<unsigned char>(std::lround(255 * r));
    rgb[1] = static_cast<unsigned char>(std::lround(255 * g));
    rgb[2] = static_cast<unsigned char>(std::lround(255 * b));
}

}  // namespace sfacade

// The resampled image of slice j2 (Graph.h:34-95): height rows of width
// values, first row at z just below +6.  A z beyond the field's last vertical
// point takes its last interval (the reference reads past the array there).
inline std::vector<stratifloat> HeatPlotData(const NodalField& U, stratifloat L1, stratifloat L3, int j2, int& width,
                                             int& height) {
    const GridParams& g = gridParams;
    const int K1 = g.N1, K3 = g.N3;
    const stratifloat zcutoff = 6;
    const int N1 = int(L1 * 50);
    const int N3 = int(2 * zcutoff * 50);
    (void)L3;  // the solver's grid holds L3 already
    std::vector<double> zd, zf;
    VerticalPointsBoth(zd, zf);
    const std::vector<double>& pts = (U.BC() == BoundaryCondition::Neumann) ? zf : zd;
    width = 2 * N1;
    height = N3;
    std::vector<stratifloat> img(size_t(width) * height);
    for (int j1 = 0; j1 < 2 * N1; ++j1) {
        const stratifloat x = j1 * L1 / N1;
        int kl = static_cast<int>(x / (L1 / K1));
        int kr = kl + 1;
        const stratifloat xl = kl * L1 / K1, xr = kr * L1 / K1;
        const stratifloat wl = (xr - x) / (xr - xl), wr = (x - xl) / (xr - xl);
        kl = ((kl % K1) + K1) % K1;
        kr = ((kr % K1) + K1) % K1;
        for (int j3 = 0; j3 < N3; ++j3) {
            const stratifloat z = j3 * 2 * zcutoff / N3 - zcutoff;
            int ka = 0, kb = 0;
            stratifloat zb = 0, za = 0;
            do {
                ++ka;
                kb = ka - 1;
                zb = stratifloat(pts[kb]);
                za = stratifloat(pts[ka]);
            } while (za < z && ka < K3 - 1);
            const stratifloat wa = (z - zb) / (za - zb), wb = (za - z) / (za - zb);
            img[size_t(N3 - 1 - j3) * width + j1] = wl * wb * U(kl, j2, kb) + wl * wa * U(kl, j2, ka) +
                                                    wr * wb * U(kr, j2, kb) + wr * wa * U(kr, j2, ka);
        }
    }
    return img;
}

// Graph.h:34-99: slice j2 of a nodal field as a PNG heat map
inline void HeatPlot(const NodalField& U, stratifloat L1, stratifloat L3, int j2, const std::string& filename) {
    int w = 0, h = 0;
    const std::vector<stratifloat> img = HeatPlotData(U, L1, L3, j2, w, h);
    const auto mm = std::minmax_element(img.begin(), img.end());
    const double lo = img.empty() ? 0 : double(*mm.first), hi = img.empty() ? 1 : double(*mm.second);
    std::vector<unsigned char> rgb(size_t(w) * h * 3);
    for (size_t k = 0; k < img.size(); ++k)
        sfacade::ramp(hi > lo ? (double(img[k]) - lo) / (hi - lo) : 0.5, &rgb[3 * k]);
    sfacade::write_png(filename, w, h, rgb);
}

// Graph.h:101-109: a modal field through its nodal values
inline void HeatPlot(const ModalField& u, stratifloat L1, stratifloat L3, int j2, const std::string& filename) {
    NodalField U(u.BC());
    u.ToNodal(U);
    HeatPlot(U, L1, L3, j2, filename);
}
