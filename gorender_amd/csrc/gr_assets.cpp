// Asset ingestion into the resident store: OBJ/MTL files as LoadObjFile reads
// them (obj.go:42-309) and PNG textures as LoadTextureFile decodes them
// (texture.go:35-63,91-103: image/png, then color.RGBAModel per pixel). Host
// code only; the arrays it produces are what gr_mesh_create and
// gr_texture_create take.

#include <zlib.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "gorender_hip.h"
#include "gr_error.h"

struct gr_obj_file {
    struct Mesh {
        std::vector<float> vertices;         // [V][4]
        std::vector<float> normals;          // [N][4]
        std::vector<int32_t> vi, ni;         // [F][3]
        std::vector<float> uvs;              // [F][3][2]
        std::vector<int32_t> tex;            // [F], index into textures or -1
    };
    struct Texture {
        gr_texture_desc desc;
        std::vector<uint8_t> rgba;
    };
    std::vector<Mesh> meshes;
    std::vector<Texture> textures;
};

namespace {

using gr_internal::fail;

// ---------------------------------------------------------------------------
// fmt.Sscanf for the formats obj.go uses
// ---------------------------------------------------------------------------

bool is_blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// Format characters: ' ' consumes zero or more blanks of the input; "%d" reads
// an optionally signed decimal int64 after blanks; "%f" reads fmt's float token
// (NaN / signed Inf / decimal or 0x-hex mantissa with exponent) after blanks and
// rounds it to float32 as strconv.ParseFloat(tok, 32) does; any other
// character must equal the next input character. Input left over after the
// format is ignored. Returns false where Sscanf returns an error.
bool go_sscanf(const std::string& line, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    const char* s = line.c_str();
    const char* e = s + line.size();
    bool ok = true;
    for (const char* f = fmt; *f && ok; f++) {
        if (*f == ' ') {
            while (s < e && is_blank(*s)) s++;
        } else if (*f == '%' && (f[1] == 'd' || f[1] == 'f')) {
            const char verb = *++f;
            while (s < e && is_blank(*s)) s++;
            const char* t0 = s;
            if (verb == 'd') {
                if (s < e && (*s == '+' || *s == '-')) s++;
                const char* d0 = s;
                while (s < e && *s >= '0' && *s <= '9') s++;
                if (s == d0 || (s < e && *s == '_')) {
                    ok = false;
                    break;
                }
                errno = 0;
                const long long v = std::strtoll(std::string(t0, s).c_str(), nullptr, 10);
                if (errno == ERANGE) ok = false;
                *va_arg(ap, int64_t*) = (int64_t)v;
            } else {
                auto accept = [&](const char* set) {
                    if (s < e && std::strchr(set, *s) && *s) {
                        s++;
                        return true;
                    }
                    return false;
                };
                std::string tok;
                if (accept("nN") && accept("aA") && accept("nN")) {
                    tok.assign(t0, s);
                } else {
                    s = t0;
                    accept("+-");
                    const char* after_sign = s;
                    if (accept("iI") && accept("nN") && accept("fF")) {
                        tok.assign(t0, s);
                    } else {
                        s = after_sign;
                        const char* digits = "0123456789";
                        const char* exp = "eE";
                        if (accept("0")) {
                            if (accept("xX")) {
                                digits = "0123456789aAbBcCdDeEfF";
                                exp = "pP";
                            }
                        }
                        while (accept(digits)) {
                        }
                        if (accept(".")) {
                            while (accept(digits)) {
                            }
                        }
                        if (accept(exp)) {
                            accept("+-");
                            while (accept("0123456789")) {
                            }
                        }
                        tok.assign(t0, s);
                    }
                }
                char* end = nullptr;
                const float v = std::strtof(tok.c_str(), &end);
                if (tok.empty() || end != tok.c_str() + tok.size()) {
                    ok = false;
                    break;
                }
                *va_arg(ap, float*) = v;
            }
        } else {
            if (s >= e || *s != *f) ok = false;
            else s++;
        }
    }
    va_end(ap);
    return ok;
}

size_t count_sub(const std::string& s, const char* sub) {
    const size_t n = std::strlen(sub);
    size_t c = 0;
    for (size_t p = s.find(sub); p != std::string::npos; p = s.find(sub, p + n)) c++;
    return c;
}

bool has_prefix(const std::string& s, const char* p) { return s.rfind(p, 0) == 0; }

// strings.TrimSpace for ASCII white space and U+0085 / U+00A0.
std::string trim_space(const std::string& s) {
    auto space_at_front = [&](size_t i, size_t& len) {
        const unsigned char c = (unsigned char)s[i];
        if (c == ' ' || (c >= '\t' && c <= '\r')) return len = 1, true;
        if (c == 0xC2 && i + 1 < s.size() &&
            ((unsigned char)s[i + 1] == 0x85 || (unsigned char)s[i + 1] == 0xA0))
            return len = 2, true;
        return false;
    };
    size_t b = 0, len = 0;
    while (b < s.size() && space_at_front(b, len)) b += len;
    size_t e = s.size();
    while (e > b) {
        const unsigned char c = (unsigned char)s[e - 1];
        if (c == ' ' || (c >= '\t' && c <= '\r')) {
            e--;
        } else if (e - b >= 2 && (c == 0x85 || c == 0xA0) && (unsigned char)s[e - 2] == 0xC2) {
            e -= 2;
        } else {
            break;
        }
    }
    return s.substr(b, e - b);
}

// bufio.Scanner lines: split on '\n', a trailing '\r' dropped; a line longer
// than bufio.MaxScanTokenSize (64 KiB) ends the scan, as the reference's
// unchecked scanner.Scan() loop does.
bool read_lines(const std::string& path, std::vector<std::string>& lines) {
    FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) return false;
    std::string data;
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, fp)) > 0) data.append(buf, n);
    std::fclose(fp);
    size_t p = 0;
    while (p < data.size()) {
        size_t q = data.find('\n', p);
        const bool last = q == std::string::npos;
        if (last) q = data.size();
        if (q - p > 65536) break;
        std::string line = data.substr(p, q - p);
        if (!line.empty() && line.back() == '\r') line.pop_back();
        lines.push_back(line);
        p = q + 1;
    }
    return true;
}

std::string path_dir(const std::string& p) {
    const size_t k = p.rfind('/');
    if (k == std::string::npos) return ".";
    if (k == 0) return "/";
    return p.substr(0, k);
}

std::string path_join(const std::string& a, const std::string& b) {
    if (a.empty() || a == ".") return b;
    return a.back() == '/' ? a + b : a + "/" + b;
}

// ---------------------------------------------------------------------------
// PNG (image/png) -> RGBA8 as color.RGBAModel converts it
// ---------------------------------------------------------------------------

uint32_t be32(const uint8_t* p) {
    return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}

// Premultiplication of non-premultiplied 8- and 16-bit colours, then >> 8
// (color.NRGBA.RGBA, color.NRGBA64.RGBA, rgbaModel in image/color).
void nrgba8(uint8_t* o, uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
    auto ch = [&](uint32_t v) { return (uint8_t)((((v | v << 8) * a) / 0xff) >> 8); };
    o[0] = ch(r);
    o[1] = ch(g);
    o[2] = ch(b);
    o[3] = (uint8_t)a;
}
void nrgba16(uint8_t* o, uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
    o[0] = (uint8_t)((r * a / 0xffff) >> 8);
    o[1] = (uint8_t)((g * a / 0xffff) >> 8);
    o[2] = (uint8_t)((b * a / 0xffff) >> 8);
    o[3] = (uint8_t)(a >> 8);
}

gr_status png_decode(const std::string& path, int32_t& W, int32_t& H, std::vector<uint8_t>& out) {
    FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) return fail(GR_ERR_INVALID_ARGUMENT, "cannot open " + path);
    std::vector<uint8_t> f;
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, fp)) > 0) f.insert(f.end(), buf, buf + n);
    std::fclose(fp);
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) {
        if (f.size() >= 3 && f[0] == 0xFF && f[1] == 0xD8 && f[2] == 0xFF)
            return fail(GR_ERR_UNSUPPORTED, "JPEG textures are not supported: " + path);
        return fail(GR_ERR_INVALID_ARGUMENT, "not a PNG file: " + path);
    }
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> pal, trns, idat;
    bool seen_ihdr = false, seen_iend = false;
    size_t p = 8;
    while (p + 12 <= f.size() && !seen_iend) {
        const uint32_t len = be32(&f[p]);
        if (len > f.size() - p - 12) return fail(GR_ERR_INVALID_ARGUMENT, "png: truncated chunk");
        const uint8_t* type = &f[p + 4];
        const uint8_t* d = &f[p + 8];
        const uint32_t crc = (uint32_t)crc32(crc32(0L, Z_NULL, 0), type, len + 4);
        if (crc != be32(d + len)) return fail(GR_ERR_INVALID_ARGUMENT, "png: invalid checksum");
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len != 13) return fail(GR_ERR_INVALID_ARGUMENT, "png: bad IHDR length");
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
            if (d[10] != 0 || d[11] != 0 || interlace > 1)
                return fail(GR_ERR_INVALID_ARGUMENT, "png: unsupported compression/filter/interlace");
            seen_ihdr = true;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            pal.assign(d, d + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            trns.assign(d, d + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            seen_iend = true;
        }
        p += 12 + (size_t)len;
    }
    if (!seen_ihdr || !seen_iend) return fail(GR_ERR_INVALID_ARGUMENT, "png: missing IHDR/IEND");
    if (w == 0 || h == 0 || w > (1u << 28) / h)
        return fail(GR_ERR_INVALID_ARGUMENT, "png: invalid or oversized dimensions");
    int chans;
    switch (ctype) {
    case 0: chans = 1; break;
    case 2: chans = 3; break;
    case 3: chans = 1; break;
    case 4: chans = 2; break;
    case 6: chans = 4; break;
    default: return fail(GR_ERR_INVALID_ARGUMENT, "png: bad color type");
    }
    const bool depth_ok =
        (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
        (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
        ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (!depth_ok) return fail(GR_ERR_INVALID_ARGUMENT, "png: bad bit depth");
    if (ctype == 3 && (pal.empty() || pal.size() % 3)) return fail(GR_ERR_INVALID_ARGUMENT, "png: bad palette");
    if (ctype == 3 && trns.size() > pal.size() / 3) return fail(GR_ERR_INVALID_ARGUMENT, "png: bad tRNS length");
    const int bpp_bits = chans * depth;
    const size_t bpp = (size_t)std::max(1, bpp_bits / 8);

    // Inflate the concatenated IDAT stream.
    z_stream zs;
    std::memset(&zs, 0, sizeof zs);
    if (inflateInit(&zs) != Z_OK) return fail(GR_ERR_OUT_OF_MEMORY, "png: inflate init");
    std::vector<uint8_t> raw;
    zs.next_in = idat.data();
    zs.avail_in = (uInt)idat.size();
    int zr;
    do {
        uint8_t chunk[1 << 16];
        zs.next_out = chunk;
        zs.avail_out = sizeof chunk;
        zr = inflate(&zs, Z_NO_FLUSH);
        if (zr != Z_OK && zr != Z_STREAM_END) {
            inflateEnd(&zs);
            return fail(GR_ERR_INVALID_ARGUMENT, "png: corrupt image data");
        }
        raw.insert(raw.end(), chunk, chunk + (sizeof chunk - zs.avail_out));
    } while (zr != Z_STREAM_END && (zs.avail_in > 0 || zs.avail_out == 0));
    inflateEnd(&zs);

    // Passes: Adam7 sub-images, or the whole image.
    struct Pass {
        uint32_t x0, y0, dx, dy;
    };
    const Pass adam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                           {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    const Pass whole[1] = {{0, 0, 1, 1}};
    const Pass* passes = interlace ? adam7 : whole;
    const int npass = interlace ? 7 : 1;

    // Decoded samples per pixel at full precision (16-bit range for every depth).
    std::vector<uint16_t> px((size_t)w * h * chans);
    std::vector<uint8_t> idx8;  // palette indexes
    if (ctype == 3) idx8.resize((size_t)w * h);
    size_t rp = 0;
    for (int ps = 0; ps < npass; ps++) {
        const Pass& P = passes[ps];
        if (P.x0 >= w || P.y0 >= h) continue;
        const uint32_t pw = (w - P.x0 + P.dx - 1) / P.dx, ph = (h - P.y0 + P.dy - 1) / P.dy;
        const size_t stride = ((size_t)pw * bpp_bits + 7) / 8;
        std::vector<uint8_t> prev(stride, 0), cur(stride);
        for (uint32_t y = 0; y < ph; y++) {
            if (rp + 1 + stride > raw.size())
                return fail(GR_ERR_INVALID_ARGUMENT, "png: not enough pixel data");
            const uint8_t ft = raw[rp];
            std::memcpy(cur.data(), &raw[rp + 1], stride);
            rp += 1 + stride;
            for (size_t i = 0; i < stride; i++) {
                const int a = i >= bpp ? cur[i - bpp] : 0, b = prev[i],
                          c = i >= bpp ? prev[i - bpp] : 0;
                int v = cur[i];
                switch (ft) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) / 2; break;
                case 4: {
                    const int pp = a + b - c, pa = std::abs(pp - a), pb = std::abs(pp - b),
                              pc = std::abs(pp - c);
                    v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                    break;
                }
                default: return fail(GR_ERR_INVALID_ARGUMENT, "png: bad filter type");
                }
                cur[i] = (uint8_t)v;
            }
            for (uint32_t x = 0; x < pw; x++) {
                const size_t o = (size_t)(P.y0 + y * P.dy) * w + (P.x0 + x * P.dx);
                for (int k = 0; k < chans; k++) {
                    uint32_t v;
                    if (depth == 16) {
                        const size_t bi = ((size_t)x * chans + k) * 2;
                        v = (uint32_t)cur[bi] << 8 | cur[bi + 1];
                    } else if (depth == 8) {
                        v = cur[(size_t)x * chans + k];
                    } else {
                        const size_t bit = (size_t)x * depth;
                        v = (cur[bit / 8] >> (8 - depth - (int)(bit % 8))) & ((1u << depth) - 1);
                    }
                    if (ctype == 3) idx8[o] = (uint8_t)v;
                    else px[o * chans + k] = (uint16_t)v;
                }
            }
            std::swap(prev, cur);
        }
    }

    out.assign((size_t)w * h * 4, 0);
    const bool has_t = !trns.empty();
    for (size_t i = 0; i < (size_t)w * h; i++) {
        uint8_t* o = &out[i * 4];
        const uint16_t* s = &px[i * chans];
        if (ctype == 3) {
            // image.Paletted: opaque RGBA entries, NRGBA where tRNS gives alpha;
            // indexes past the palette read opaque black.
            const uint32_t k = idx8[i];
            if (k * 3 + 2 >= pal.size()) {
                o[0] = o[1] = o[2] = 0;
                o[3] = 0xff;
            } else if (k < trns.size()) {
                nrgba8(o, pal[3 * k], pal[3 * k + 1], pal[3 * k + 2], trns[k]);
            } else {
                o[0] = pal[3 * k];
                o[1] = pal[3 * k + 1];
                o[2] = pal[3 * k + 2];
                o[3] = 0xff;
            }
        } else if (ctype == 0) {
            // Gray, Gray16; low depths scale to 8 bits; tRNS makes the matching
            // sample transparent.
            uint32_t y = s[0];
            const bool transparent = has_t && trns.size() >= 2 && y == ((uint32_t)trns[0] << 8 | trns[1]);
            if (depth < 8) y = y * (255u / ((1u << depth) - 1));
            if (depth == 16) {
                if (transparent) std::memset(o, 0, 4);
                else o[0] = o[1] = o[2] = (uint8_t)(y >> 8), o[3] = 0xff;
            } else {
                if (transparent) std::memset(o, 0, 4);
                else o[0] = o[1] = o[2] = (uint8_t)y, o[3] = 0xff;
            }
        } else if (ctype == 2) {
            const bool transparent =
                has_t && trns.size() >= 6 && s[0] == ((uint32_t)trns[0] << 8 | trns[1]) &&
                s[1] == ((uint32_t)trns[2] << 8 | trns[3]) && s[2] == ((uint32_t)trns[4] << 8 | trns[5]);
            if (transparent) {
                std::memset(o, 0, 4);
            } else if (depth == 16) {
                o[0] = (uint8_t)(s[0] >> 8);
                o[1] = (uint8_t)(s[1] >> 8);
                o[2] = (uint8_t)(s[2] >> 8);
                o[3] = 0xff;
            } else {
                o[0] = (uint8_t)s[0];
                o[1] = (uint8_t)s[1];
                o[2] = (uint8_t)s[2];
                o[3] = 0xff;
            }
        } else if (ctype == 4) {
            if (depth == 16) nrgba16(o, s[0], s[0], s[0], s[1]);
            else nrgba8(o, s[0], s[0], s[0], s[1]);
        } else {
            if (depth == 16) nrgba16(o, s[0], s[1], s[2], s[3]);
            else nrgba8(o, s[0], s[1], s[2], s[3]);
        }
    }
    W = (int32_t)w;
    H = (int32_t)h;
    return GR_OK;
}

// ---------------------------------------------------------------------------
// OBJ / MTL (obj.go)
// ---------------------------------------------------------------------------

struct Material {
    std::string name, map_kd;
};

// parseMtlLibFile (obj.go:156-193).
gr_status parse_mtl(const std::string& path, std::vector<Material>& mats) {
    std::vector<std::string> lines;
    if (!read_lines(path, lines)) return fail(GR_ERR_INVALID_ARGUMENT, "failed to parse material library: cannot open " + path);
    bool have = false;
    Material cur;
    for (const std::string& raw : lines) {
        const std::string line = trim_space(raw);
        if (line.empty()) continue;
        if (has_prefix(line, "newmtl ")) {
            if (have) mats.push_back(cur);
            cur = Material{line.substr(7), ""};
            have = true;
        } else if (has_prefix(line, "map_Kd ")) {
            // The reference dereferences a nil material here (a panic).
            if (!have) return fail(GR_ERR_INVALID_ARGUMENT, "map_Kd before newmtl in " + path);
            cur.map_kd = line.substr(7);
        }
    }
    if (have) mats.push_back(cur);
    return GR_OK;
}

gr_status load_obj(const std::string& filename, bool single_mesh, gr_obj_file& out) {
    std::vector<std::string> lines;
    if (!read_lines(filename, lines)) return fail(GR_ERR_INVALID_ARGUMENT, "cannot open " + filename);
    const std::string dirname = path_dir(filename);

    gr_obj_file::Mesh cur;             // c.Vertices, c.Faces, c.VertexNormals
    std::vector<float> tverts;         // c.TextureVertices
    int64_t voff = 0, tvoff = 0, vnoff = 0;
    std::map<std::string, int32_t> material_tex;  // c.Textures
    std::map<std::string, int32_t> file_tex;      // textureFiles
    int32_t default_tex = -1;
    int32_t current_tex = -1;

    auto finish_mesh = [&]() { out.meshes.push_back(std::move(cur)); cur = gr_obj_file::Mesh(); };
    auto clear = [&]() {  // ObjContext.Clear (obj.go:31-40)
        voff += (int64_t)(cur.vertices.size() / 4);
        tvoff += (int64_t)(tverts.size() / 2);
        vnoff += (int64_t)(cur.normals.size() / 4);
        tverts.clear();
    };

    for (const std::string& raw : lines) {
        const std::string line = trim_space(raw);
        if (line.empty()) continue;
        if (has_prefix(line, "mtllib ")) {
            std::vector<Material> mats;
            gr_status st = parse_mtl(path_join(dirname, line.substr(7)), mats);
            if (st) return st;
            for (const Material& m : mats) {
                if (m.map_kd.empty()) {
                    if (default_tex < 0) {
                        // &Texture{color: {255,0,255,255}}: a solid colour texture.
                        gr_obj_file::Texture t;
                        std::memset(&t.desc, 0, sizeof t.desc);
                        t.desc.type = GR_TEXTURE_SOLID;
                        t.desc.scale = 0.0f;
                        const uint8_t magenta[4] = {255, 0, 255, 255};
                        std::memcpy(t.desc.color, magenta, 4);
                        out.textures.push_back(std::move(t));
                        default_tex = (int32_t)out.textures.size() - 1;
                    }
                    material_tex[m.name] = default_tex;
                } else if (file_tex.count(m.map_kd)) {
                    material_tex[m.name] = file_tex[m.map_kd];
                } else {
                    const std::string tp = m.map_kd[0] == '/' ? m.map_kd : path_join(dirname, m.map_kd);
                    gr_obj_file::Texture t;
                    int32_t w, h;
                    gr_status ts = png_decode(tp, w, h, t.rgba);
                    if (ts) return ts;
                    std::memset(&t.desc, 0, sizeof t.desc);
                    t.desc.type = gr_texture_kind_for_size(w, h);
                    t.desc.width = w;
                    t.desc.height = h;
                    t.desc.scale = 1.0f;
                    out.textures.push_back(std::move(t));
                    file_tex[m.map_kd] = material_tex[m.name] = (int32_t)out.textures.size() - 1;
                }
            }
        } else if (has_prefix(line, "o ")) {
            if (!cur.vertices.empty() && !single_mesh) {
                clear();
                finish_mesh();
            }
        } else if (has_prefix(line, "v ")) {
            float x, y, z;
            if (!go_sscanf(line, "v %f %f %f", &x, &y, &z))
                return fail(GR_ERR_INVALID_ARGUMENT, "bad vertex line: " + line);
            cur.vertices.insert(cur.vertices.end(), {x, y, z, 1.0f});
        } else if (has_prefix(line, "vt ")) {
            float u, v;
            if (!go_sscanf(line, "vt %f %f", &u, &v))
                return fail(GR_ERR_INVALID_ARGUMENT, "bad texture vertex line: " + line);
            tverts.insert(tverts.end(), {u, v});
        } else if (has_prefix(line, "vn ")) {
            float x, y, z;
            if (!go_sscanf(line, "vn %f %f %f", &x, &y, &z))
                return fail(GR_ERR_INVALID_ARGUMENT, "bad vertex normal line: " + line);
            cur.normals.insert(cur.normals.end(), {x, y, z, 1.0f});
        } else if (has_prefix(line, "usemtl ")) {
            auto it = material_tex.find(line.substr(7));
            current_tex = it == material_tex.end() ? -1 : it->second;
        } else if (has_prefix(line, "f ")) {
            // parseFace (obj.go:60-151)
            if (count_sub(line, " ") != 3)
                return fail(GR_ERR_INVALID_ARGUMENT, "mesh is not triangulated");
            int64_t v[3] = {0, 0, 0}, t[3] = {0, 0, 0}, nn[3] = {0, 0, 0};
            bool ok, has_t = false;
            if (count_sub(line, "//") == 3) {
                // The reference scans the third normal into vn1 (obj.go:77-82):
                // the second normal index is the third one's and the third is
                // left at 0, i.e. -1 after the offset.
                int64_t n1b = 0;
                ok = go_sscanf(line, "f %d//%d %d//%d %d//%d", &v[0], &nn[0], &v[1], &nn[1], &v[2], &n1b);
                nn[1] = n1b;
                nn[2] = 0;
            } else if (count_sub(line, "/") == 3) {
                ok = go_sscanf(line, "f %d/%d %d/%d %d/%d", &v[0], &t[0], &v[1], &t[1], &v[2], &t[2]);
                has_t = true;
            } else if (count_sub(line, "/") == 6) {
                ok = go_sscanf(line, "f %d/%d/%d %d/%d/%d %d/%d/%d", &v[0], &t[0], &nn[0], &v[1],
                               &t[1], &nn[1], &v[2], &t[2], &nn[2]);
                has_t = true;
            } else {
                ok = go_sscanf(line, "f %d %d %d", &v[0], &v[1], &v[2]);
            }
            const bool has_n = count_sub(line, "//") == 3 || count_sub(line, "/") == 6;
            float uv[6] = {0, 0, 0, 0, 0, 0};
            if (has_t) {
                // UVs are looked up while parsing (a panic when out of range).
                const int64_t ntv = (int64_t)(tverts.size() / 2);
                for (int k = 0; k < 3; k++) {
                    const int64_t ti = t[k] - tvoff - 1;
                    if (ti < 0 || ti >= ntv)
                        return fail(GR_ERR_INDEX_OUT_OF_RANGE, "texture vertex index out of range");
                    uv[2 * k] = tverts[2 * ti];
                    uv[2 * k + 1] = tverts[2 * ti + 1];
                }
            }
            if (!ok) return fail(GR_ERR_INVALID_ARGUMENT, "bad face line: " + line);
            for (int k = 0; k < 3; k++) {
                const int64_t vi = v[k] - voff - 1;
                const int64_t ni = has_n ? nn[k] - vnoff - 1 : 0;
                if (vi < INT32_MIN || vi > INT32_MAX || ni < INT32_MIN || ni > INT32_MAX)
                    return fail(GR_ERR_INDEX_OUT_OF_RANGE, "face index out of range");
                cur.vi.push_back((int32_t)vi);
                cur.ni.push_back((int32_t)ni);
            }
            cur.uvs.insert(cur.uvs.end(), uv, uv + 6);
            cur.tex.push_back(current_tex);
        }
    }
    if (!cur.vertices.empty()) finish_mesh();
    if (out.meshes.empty()) return fail(GR_ERR_INVALID_ARGUMENT, "obj file does not have any vertices data");
    return GR_OK;
}

}  // namespace

extern "C" {

gr_status gr_obj_open(const char* path, int32_t single_mesh, gr_obj_file** out) {
    try {
        if (!path || !out) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        std::unique_ptr<gr_obj_file> f(new gr_obj_file());
        gr_status st = load_obj(path, single_mesh != 0, *f);
        if (st) return st;
        *out = f.release();
        return GR_OK;
    } catch (...) {
        return fail(GR_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

void gr_obj_close(gr_obj_file* obj) { delete obj; }

int32_t gr_obj_num_meshes(const gr_obj_file* obj) { return obj ? (int32_t)obj->meshes.size() : 0; }

int32_t gr_obj_num_textures(const gr_obj_file* obj) {
    return obj ? (int32_t)obj->textures.size() : 0;
}

gr_status gr_obj_mesh(const gr_obj_file* obj, int32_t i, gr_mesh_desc* d) {
    if (!obj || !d) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    if (i < 0 || i >= (int32_t)obj->meshes.size()) return fail(GR_ERR_INVALID_HANDLE, "no such mesh");
    const gr_obj_file::Mesh& m = obj->meshes[i];
    std::memset(d, 0, sizeof *d);
    d->num_vertices = (int32_t)(m.vertices.size() / 4);
    d->vertices = m.vertices.data();
    d->num_vertex_normals = (int32_t)(m.normals.size() / 4);
    d->vertex_normals = m.normals.empty() ? nullptr : m.normals.data();
    d->num_faces = (int32_t)m.tex.size();
    d->vertex_indices = m.vi.data();
    d->normal_indices = m.ni.data();
    d->uvs = m.uvs.data();
    d->face_textures = m.tex.data();
    d->face_normals = nullptr;
    return GR_OK;
}

gr_status gr_obj_texture(const gr_obj_file* obj, int32_t i, gr_texture_desc* d,
                         const uint8_t** rgba) {
    if (!obj || !d) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    if (i < 0 || i >= (int32_t)obj->textures.size())
        return fail(GR_ERR_INVALID_HANDLE, "no such texture");
    *d = obj->textures[i].desc;
    if (rgba) *rgba = obj->textures[i].rgba.empty() ? nullptr : obj->textures[i].rgba.data();
    return GR_OK;
}

gr_status gr_obj_upload(gr_context* ctx, const gr_obj_file* obj, int32_t* mesh_handles,
                        int32_t max_meshes, int32_t* num_meshes) {
    if (!ctx || !obj) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    const int32_t nm = (int32_t)obj->meshes.size();
    if (num_meshes) *num_meshes = nm;
    if (nm > max_meshes || (nm && !mesh_handles))
        return fail(GR_ERR_CAPACITY, "mesh_handles holds fewer entries than the file's meshes");
    try {
        std::vector<int32_t> th(obj->textures.size());
        for (size_t k = 0; k < obj->textures.size(); k++) {
            const auto& t = obj->textures[k];
            gr_status st = gr_texture_create(ctx, &t.desc, t.rgba.empty() ? nullptr : t.rgba.data(), &th[k]);
            if (st) return st;
        }
        for (int32_t i = 0; i < nm; i++) {
            const gr_obj_file::Mesh& m = obj->meshes[i];
            std::vector<int32_t> ft(m.tex.size());
            for (size_t k = 0; k < ft.size(); k++) ft[k] = m.tex[k] < 0 ? -1 : th[m.tex[k]];
            gr_mesh_desc d;
            gr_obj_mesh(obj, i, &d);
            d.face_textures = ft.data();
            gr_status st = gr_mesh_create(ctx, &d, &mesh_handles[i]);
            if (st) return st;
        }
        return GR_OK;
    } catch (...) {
        return fail(GR_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

gr_status gr_png_decode(const char* path, int32_t* width, int32_t* height, uint8_t** rgba) {
    if (!path || !width || !height || !rgba) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    *rgba = nullptr;
    try {
        std::vector<uint8_t> px;
        gr_status st = png_decode(path, *width, *height, px);
        if (st) return st;
        *rgba = (uint8_t*)std::malloc(px.size());
        if (!*rgba) return fail(GR_ERR_OUT_OF_MEMORY, "host allocation failed");
        std::memcpy(*rgba, px.data(), px.size());
        return GR_OK;
    } catch (...) {
        return fail(GR_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

void gr_host_free(void* p) { std::free(p); }

gr_status gr_texture_load(gr_context* ctx, const char* path, int32_t* out_handle) {
    if (!ctx || !path || !out_handle) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    try {
        std::vector<uint8_t> px;
        int32_t w, h;
        gr_status st = png_decode(path, w, h, px);
        if (st) return st;
        gr_texture_desc d;
        std::memset(&d, 0, sizeof d);
        d.type = gr_texture_kind_for_size(w, h);
        d.width = w;
        d.height = h;
        d.scale = 1.0f;
        return gr_texture_create(ctx, &d, px.data(), out_handle);
    } catch (...) {
        return fail(GR_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

}  // extern "C"
