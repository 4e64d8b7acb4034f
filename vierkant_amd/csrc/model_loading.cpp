// model_loading.cpp - scene files into the path tracer's flat scene: Wavefront OBJ (+ MTL) and glTF 2.0
// (.gltf with external or embedded buffers, .glb). Host code only; nothing here touches the GPU.
//
// The reference loads models through tinyobjloader and tinygltf (src/model/wavefront_obj.cpp:11-137,
// src/model/gltf.cpp:136-660, 933-1218); neither library is in the checkout, so both formats are parsed here
// from their published specifications, and the reference's own mapping from the parsed data to vierkant's
// geometry and material_t is restated:
//   OBJ   every face corner becomes its own vertex (create_geometry, wavefront_obj.cpp:13-45); positions and
//         texture coordinates by index, normals only when the file has any; without normals, vertex normals
//         are computed as Geometry::compute_vertex_normals does (src/Geometry.cpp:106-137); a shape takes the
//         material of its first face (:128); materials map as :84-99 (alpha = clamp(1 - dissolve), roughness =
//         clamp(max(Pr, (1 - Ns)^2)), metalness = Pm, ior = Ni, clearcoat roughness = Pcr); a file without
//         materials gets one material_t with its defaults.
//   glTF  indices u8/u16/u32 or 0..n-1 (create_geometry, gltf.cpp:198-312); POSITION, NORMAL, TEXCOORD_0 as
//         float; missing normals computed as above; materials as convert_material (:313-660): base color,
//         emissive factor, metallic and roughness factors, doubleSided, alphaMode/alphaCutoff, and the
//         KHR_materials_emissive_strength / transmission / ior / clearcoat / sheen / diffuse_transmission /
//         volume (attenuation)
//         factors. Node transforms (matrix, or translation * rotation * scale, :136-160) are composed from the
//         scene's roots and baked into world-space vertices in double precision, rounded to float once.
// Base color textures are decoded (PNG, via zlib's inflate) into RGBA8; other images are counted as skipped.
#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <zlib.h>

#include "vkrt.h"

namespace
{

thread_local std::string g_model_error;

struct load_error : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------------------------------------
// JSON (RFC 8259): objects keep their keys sorted, numbers are doubles
// ---------------------------------------------------------------------------------------------
struct json
{
    enum kind_t { null_t, bool_t, number_t, string_t, array_t, object_t } kind = null_t;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<json> arr;
    std::map<std::string, json> obj;

    bool has(const std::string &k) const { return kind == object_t && obj.count(k); }
    const json &operator[](const std::string &k) const
    {
        static const json empty;
        if(kind != object_t) return empty;
        auto it = obj.find(k);
        return it == obj.end() ? empty : it->second;
    }
    const json &operator[](size_t i) const
    {
        static const json empty;
        return (kind == array_t && i < arr.size()) ? arr[i] : empty;
    }
    size_t size() const { return kind == array_t ? arr.size() : 0; }
    double number(double def) const { return kind == number_t ? num : def; }
    //! the number when it is an integer that fits an int, else `def` (NaN, fractions and out-of-range values
    //! never reach the conversion)
    int integer(int def) const
    {
        if(kind != number_t || !(num >= -2147483648.0 && num <= 2147483647.0) || num != std::floor(num)) return def;
        return static_cast<int>(num);
    }
};

struct json_parser
{
    const char *p, *end;

    [[noreturn]] void fail(const char *what) { throw load_error(std::string("JSON: ") + what); }
    void ws()
    {
        while(p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool lit(const char *s)
    {
        const size_t n = std::strlen(s);
        if((size_t)(end - p) >= n && std::memcmp(p, s, n) == 0)
        {
            p += n;
            return true;
        }
        return false;
    }
    static void put_utf8(std::string &out, uint32_t cp)
    {
        if(cp < 0x80) out += static_cast<char>(cp);
        else if(cp < 0x800)
        {
            out += static_cast<char>(0xC0 | (cp >> 6));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        }
        else if(cp < 0x10000)
        {
            out += static_cast<char>(0xE0 | (cp >> 12));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        }
        else
        {
            out += static_cast<char>(0xF0 | (cp >> 18));
            out += static_cast<char>(0x80 | ((cp >> 12) & 0x3F));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        }
    }
    uint32_t hex4()
    {
        if(end - p < 4) fail("short \\u escape");
        uint32_t v = 0;
        for(int k = 0; k < 4; ++k)
        {
            const char c = *p++;
            v <<= 4;
            if(c >= '0' && c <= '9') v |= c - '0';
            else if(c >= 'a' && c <= 'f') v |= c - 'a' + 10;
            else if(c >= 'A' && c <= 'F') v |= c - 'A' + 10;
            else fail("bad \\u escape");
        }
        return v;
    }
    std::string string()
    {
        if(p >= end || *p != '"') fail("expected a string");
        ++p;
        std::string out;
        while(true)
        {
            if(p >= end) fail("unterminated string");
            const char c = *p++;
            if(c == '"') break;
            if(c != '\\')
            {
                out += c;
                continue;
            }
            if(p >= end) fail("unterminated escape");
            const char e = *p++;
            switch(e)
            {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u':
                {
                    uint32_t cp = hex4();
                    if(cp >= 0xD800 && cp < 0xDC00 && lit("\\u"))
                    {
                        const uint32_t lo = hex4();
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(out, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
        return out;
    }
    json value(int depth = 0)
    {
        if(depth > 256) fail("nesting too deep");
        ws();
        if(p >= end) fail("unexpected end");
        json v;
        if(*p == '{')
        {
            ++p;
            v.kind = json::object_t;
            ws();
            if(p < end && *p == '}')
            {
                ++p;
                return v;
            }
            while(true)
            {
                ws();
                std::string k = string();
                ws();
                if(p >= end || *p != ':') fail("expected ':'");
                ++p;
                v.obj[k] = value(depth + 1);
                ws();
                if(p < end && *p == ',')
                {
                    ++p;
                    continue;
                }
                if(p < end && *p == '}')
                {
                    ++p;
                    break;
                }
                fail("expected ',' or '}'");
            }
        }
        else if(*p == '[')
        {
            ++p;
            v.kind = json::array_t;
            ws();
            if(p < end && *p == ']')
            {
                ++p;
                return v;
            }
            while(true)
            {
                v.arr.push_back(value(depth + 1));
                ws();
                if(p < end && *p == ',')
                {
                    ++p;
                    continue;
                }
                if(p < end && *p == ']')
                {
                    ++p;
                    break;
                }
                fail("expected ',' or ']'");
            }
        }
        else if(*p == '"')
        {
            v.kind = json::string_t;
            v.str = string();
        }
        else if(lit("true")) v.kind = json::bool_t, v.b = true;
        else if(lit("false")) v.kind = json::bool_t, v.b = false;
        else if(lit("null")) v.kind = json::null_t;
        else
        {
            std::string num;
            while(p < end && (std::strchr("+-0123456789.eE", *p) != nullptr)) num += *p++;
            if(num.empty()) fail("unexpected character");
            char *e = nullptr;
            v.kind = json::number_t;
            v.num = std::strtod(num.c_str(), &e);
            if(!e || *e) fail("bad number");
        }
        return v;
    }
};

json parse_json(const char *data, size_t n)
{
    json_parser jp{data, data + n};
    json v = jp.value();
    jp.ws();
    if(jp.p != jp.end) jp.fail("trailing characters");
    return v;
}

// ---------------------------------------------------------------------------------------------
// shared helpers
// ---------------------------------------------------------------------------------------------
std::string read_file(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if(!f) throw load_error("cannot open " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

std::string directory_of(const std::string &path)
{
    const size_t k = path.find_last_of('/');
    return k == std::string::npos ? std::string() : path.substr(0, k + 1);
}

std::string lower_extension(const std::string &path)
{
    const size_t k = path.find_last_of('.');
    std::string e = k == std::string::npos ? std::string() : path.substr(k + 1);
    for(char &c: e) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return e;
}

struct v3f
{
    float x, y, z;
};

//! glm::normalize: v * inversesqrt(dot(v, v)), here with a correctly rounded 1 / sqrt
v3f normalize(v3f v)
{
    const float d = (v.x * v.x + v.y * v.y) + v.z * v.z;
    const float s = 1.0f / std::sqrt(d);
    return v3f{v.x * s, v.y * s, v.z * s};
}

//! Geometry::compute_vertex_normals: normalized face normals summed per vertex in index order, normalized
void compute_vertex_normals(const std::vector<float> &pos, const std::vector<uint32_t> &idx, std::vector<float> &nrm)
{
    nrm.assign(pos.size(), 0.f);
    if(idx.size() < 3) return;
    for(size_t i = 0; i + 2 < idx.size(); i += 3)
    {
        const uint32_t a = idx[i], b = idx[i + 1], c = idx[i + 2];
        const v3f A{pos[3 * a], pos[3 * a + 1], pos[3 * a + 2]}, B{pos[3 * b], pos[3 * b + 1], pos[3 * b + 2]},
                C{pos[3 * c], pos[3 * c + 1], pos[3 * c + 2]};
        const v3f x{B.x - A.x, B.y - A.y, B.z - A.z}, y{C.x - A.x, C.y - A.y, C.z - A.z};
        // glm::cross
        const v3f cr{x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
        const v3f n = normalize(cr);
        for(uint32_t v: {a, b, c})
        {
            nrm[3 * v] += n.x;
            nrm[3 * v + 1] += n.y;
            nrm[3 * v + 2] += n.z;
        }
    }
    for(size_t v = 0; v < pos.size() / 3; ++v)
    {
        const v3f n = normalize(v3f{nrm[3 * v], nrm[3 * v + 1], nrm[3 * v + 2]});
        nrm[3 * v] = n.x;
        nrm[3 * v + 1] = n.y;
        nrm[3 * v + 2] = n.z;
    }
}

//! Geometry::compute_tangents (src/Geometry.cpp:196-254): per triangle the texture-space s and t directions,
//! summed per vertex in index order; then Gram-Schmidt against the normal, the sign from the handedness test
//! as the reference writes it (flipped when dot(cross(n, t), b) is not negative). Degenerate coordinates give
//! non-finite tangents, as there.
void compute_tangents(const std::vector<float> &pos, const std::vector<float> &nrm, const std::vector<float> &uv,
                      const std::vector<uint32_t> &idx, std::vector<float> &tan)
{
    const size_t nv = pos.size() / 3;
    tan.assign(3 * nv, 0.f);
    if(idx.size() % 3 || uv.size() != 2 * nv || nrm.size() != 3 * nv) return;
    std::vector<float> ts(3 * nv, 0.f), bs(3 * nv, 0.f);
    for(size_t i = 0; i < idx.size(); i += 3)
    {
        const uint32_t a = idx[i], b = idx[i + 1], c = idx[i + 2];
        const float x1 = pos[3 * b] - pos[3 * a], x2 = pos[3 * c] - pos[3 * a];
        const float y1 = pos[3 * b + 1] - pos[3 * a + 1], y2 = pos[3 * c + 1] - pos[3 * a + 1];
        const float z1 = pos[3 * b + 2] - pos[3 * a + 2], z2 = pos[3 * c + 2] - pos[3 * a + 2];
        const float s1 = uv[2 * b] - uv[2 * a], s2 = uv[2 * c] - uv[2 * a];
        const float t1 = uv[2 * b + 1] - uv[2 * a + 1], t2 = uv[2 * c + 1] - uv[2 * a + 1];
        const float r = 1.f / (s1 * t2 - s2 * t1);
        const float sd[3] = {(t2 * x1 - t1 * x2) * r, (t2 * y1 - t1 * y2) * r, (t2 * z1 - t1 * z2) * r};
        const float td[3] = {(s1 * x2 - s2 * x1) * r, (s1 * y2 - s2 * y1) * r, (s1 * z2 - s2 * z1) * r};
        for(uint32_t v: {a, b, c})
            for(int k = 0; k < 3; ++k)
            {
                ts[3 * v + k] += sd[k];
                bs[3 * v + k] += td[k];
            }
    }
    for(size_t v = 0; v < nv; ++v)
    {
        const v3f n{nrm[3 * v], nrm[3 * v + 1], nrm[3 * v + 2]}, t{ts[3 * v], ts[3 * v + 1], ts[3 * v + 2]};
        const v3f bt{bs[3 * v], bs[3 * v + 1], bs[3 * v + 2]};
        const float nt = (n.x * t.x + n.y * t.y) + n.z * t.z;
        const v3f o = normalize(v3f{t.x - n.x * nt, t.y - n.y * nt, t.z - n.z * nt});
        const v3f c{n.y * t.z - t.y * n.z, n.z * t.x - t.z * n.x, n.x * t.y - t.x * n.y};
        const float sign = ((c.x * bt.x + c.y * bt.y) + c.z * bt.z) < 0.f ? 1.f : -1.f;
        tan[3 * v] = o.x * sign;
        tan[3 * v + 1] = o.y * sign;
        tan[3 * v + 2] = o.z * sign;
    }
}

//! vierkant::material_t defaults (include/vierkant/Material.hpp:85-142) in the tracer's record
vkrt_material default_material()
{
    vkrt_material m{};
    for(int k = 0; k < 4; ++k) m.color[k] = 1.f;
    m.emission[3] = 1.f;// emissive_strength
    m.roughness = 1.f;
    m.metalness = 0.f;
    m.ior = 1.5f;
    for(int k = 0; k < 3; ++k) m.diffuse_transmission_color[k] = 1.f;
    return m;
}

//! an RGBA8 image, rows from the top
struct image_rgba8
{
    uint32_t width = 0, height = 0;
    std::vector<uint8_t> rgba;
};

uint32_t be32(const uint8_t *p) { return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3]; }

//! PNG (ISO/IEC 15948): 1-16 bit depths of every color type, not interlaced; 16-bit samples keep their high
//! byte, gray expands to rgb, palettes and tRNS to rgba
bool decode_png(const uint8_t *data, size_t size, image_rgba8 &out, std::string &why)
{
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    if(size < 8 || std::memcmp(data, sig, 8) != 0)
    {
        why = "not a PNG";
        return false;
    }
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    size_t off = 8;
    while(off + 12 <= size)
    {
        const uint32_t len = be32(data + off);
        const uint8_t *type = data + off + 4, *body = data + off + 8;
        if(off + 12 + (size_t)len > size)
        {
            why = "truncated chunk";
            return false;
        }
        if(!std::memcmp(type, "IHDR", 4) && len >= 13)
        {
            w = be32(body);
            h = be32(body + 4);
            depth = body[8];
            ctype = body[9];
            interlace = body[12];
        }
        else if(!std::memcmp(type, "PLTE", 4)) plte.assign(body, body + len);
        else if(!std::memcmp(type, "tRNS", 4)) trns.assign(body, body + len);
        else if(!std::memcmp(type, "IDAT", 4)) idat.insert(idat.end(), body, body + len);
        else if(!std::memcmp(type, "IEND", 4)) break;
        off += 12 + (size_t)len;
    }
    const int channels = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    if(!w || !h || !channels || interlace > 1 || !(depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16) ||
       (ctype != 0 && ctype != 3 && depth < 8) || (ctype == 3 && depth > 8) || (uint64_t)w * h > (1u << 28))
    {
        why = "unsupported PNG header";
        return false;
    }
    const size_t bits = (size_t)channels * depth, bpp = std::max<size_t>(1, bits / 8);
    // the sub-images of the stream: the whole image, or Adam7's seven passes (start and step per axis)
    struct pass_t
    {
        uint32_t x0, y0, dx, dy;
    };
    static const pass_t adam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                    {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    const pass_t whole{0, 0, 1, 1};
    const int n_pass = interlace ? 7 : 1;
    auto pass_size = [&](const pass_t &q, uint32_t &pw, uint32_t &ph) {
        pw = w > q.x0 ? (w - q.x0 + q.dx - 1) / q.dx : 0;
        ph = h > q.y0 ? (h - q.y0 + q.dy - 1) / q.dy : 0;
    };
    size_t total = 0;
    for(int k = 0; k < n_pass; ++k)
    {
        uint32_t pw, ph;
        pass_size(interlace ? adam7[k] : whole, pw, ph);
        if(pw && ph) total += (((size_t)pw * bits + 7) / 8 + 1) * ph;
    }
    std::vector<uint8_t> raw(total);
    z_stream zs{};
    if(inflateInit(&zs) != Z_OK)
    {
        why = "zlib";
        return false;
    }
    zs.next_in = idat.data();
    zs.avail_in = static_cast<uInt>(idat.size());
    zs.next_out = raw.data();
    zs.avail_out = static_cast<uInt>(raw.size());
    const int zr = inflate(&zs, Z_FINISH);
    inflateEnd(&zs);
    if(zr != Z_STREAM_END || zs.avail_out != 0)
    {
        why = "bad image data";
        return false;
    }
    // samples at their stored depth, channels interleaved, for the whole image
    std::vector<uint16_t> px((size_t)w * h * channels);
    const uint8_t *in = raw.data();
    for(int k = 0; k < n_pass; ++k)
    {
        const pass_t q = interlace ? adam7[k] : whole;
        uint32_t pw, ph;
        pass_size(q, pw, ph);
        if(!pw || !ph) continue;
        const size_t row = ((size_t)pw * bits + 7) / 8;
        // undo the per-row filters (none, sub, up, average, Paeth)
        std::vector<uint8_t> cur(row), prev(row, 0);
        for(uint32_t y = 0; y < ph; ++y, in += row + 1)
        {
            const uint8_t f = in[0];
            const uint8_t *src = in + 1;
            for(size_t x = 0; x < row; ++x)
            {
                const int a = x >= bpp ? cur[x - bpp] : 0, b = prev[x], c = x >= bpp ? prev[x - bpp] : 0;
                int v = src[x];
                switch(f)
                {
                    case 0: break;
                    case 1: v += a; break;
                    case 2: v += b; break;
                    case 3: v += (a + b) / 2; break;
                    case 4:
                    {
                        const int pp = a + b - c, pa = std::abs(pp - a), pb = std::abs(pp - b), pc = std::abs(pp - c);
                        v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                        break;
                    }
                    default: why = "bad filter"; return false;
                }
                cur[x] = static_cast<uint8_t>(v);
            }
            for(uint32_t x = 0; x < pw; ++x)
                for(int ch = 0; ch < channels; ++ch)
                {
                    const size_t i = (size_t)x * channels + ch;
                    uint32_t v;
                    if(depth == 16) v = (uint32_t)cur[2 * i] << 8 | cur[2 * i + 1];
                    else if(depth == 8) v = cur[i];
                    else v = (cur[i * depth / 8] >> (8 - depth - (i * depth) % 8)) & ((1u << depth) - 1u);
                    px[(((size_t)(q.y0 + y * q.dy)) * w + q.x0 + x * q.dx) * channels + ch] = static_cast<uint16_t>(v);
                }
            std::swap(cur, prev);
        }
    }
    // to 8 bits: 16-bit samples keep their high byte, grey levels below 8 bits scale to 0..255
    auto to8 = [&](uint32_t v) -> uint8_t {
        return static_cast<uint8_t>(depth == 16 ? v >> 8 : depth == 8 ? v : v * 255u / ((1u << depth) - 1u));
    };
    // tRNS: palette alphas, or the one grey / RGB value (at the stored depth) that is fully transparent
    const bool key = (ctype == 0 && trns.size() >= 2) || (ctype == 2 && trns.size() >= 6);
    uint32_t kv[3] = {0, 0, 0};
    if(key)
        for(int k = 0; k < (ctype == 0 ? 1 : 3); ++k) kv[k] = ((uint32_t)trns[2 * k] << 8 | trns[2 * k + 1]) & ((1u << depth) - 1u);
    out.width = w;
    out.height = h;
    out.rgba.resize((size_t)w * h * 4);
    for(size_t i = 0; i < (size_t)w * h; ++i)
    {
        const uint16_t *s = px.data() + i * channels;
        uint8_t *o = out.rgba.data() + i * 4;
        if(ctype == 3)
        {
            if(3 * (size_t)s[0] + 2 >= plte.size())
            {
                why = "palette index out of range";
                return false;
            }
            o[0] = plte[3 * s[0]];
            o[1] = plte[3 * s[0] + 1];
            o[2] = plte[3 * s[0] + 2];
            o[3] = s[0] < trns.size() ? trns[s[0]] : 255;
        }
        else if(ctype == 0 || ctype == 4)
        {
            o[0] = o[1] = o[2] = to8(s[0]);
            o[3] = ctype == 4 ? to8(s[1]) : (key && s[0] == kv[0] ? 0 : 255);
        }
        else
        {
            for(int k = 0; k < 3; ++k) o[k] = to8(s[k]);
            o[3] = ctype == 6 ? to8(s[3]) : (key && s[0] == kv[0] && s[1] == kv[1] && s[2] == kv[2] ? 0 : 255);
        }
    }
    return true;
}

}// namespace

//! the loaded scene: world-space vertices, triangles grouped into objects (one per OBJ shape / glTF primitive
//! instance) with one material each
struct vkrt_model
{
    std::vector<float> positions, normals, texcoords, tangents;
    std::vector<uint32_t> indices, tri_object, object_material;
    std::vector<vkrt_material> materials;
    std::vector<std::string> object_names, material_names;
    std::vector<image_rgba8> images;
    //! (image, sampler) pairs: one texture per distinct pair, as the reference keys textures by
    //! (texture id, sampler id)
    std::vector<std::array<uint32_t, 5>> texture_refs;
    std::vector<vkrt_texture> textures;// views of `images`, filled when the model is complete
    uint32_t n_objects = 0, n_skipped_textures = 0;

    //! decode one image; returns 1 + its image index, or 0 (counted as skipped) when it cannot be decoded
    uint32_t add_image(const uint8_t *data, size_t size)
    {
        image_rgba8 img;
        std::string why;
        if(!decode_png(data, size, img, why))
        {
            ++n_skipped_textures;
            return 0;
        }
        images.push_back(std::move(img));
        return static_cast<uint32_t>(images.size());
    }
    //! 1 + the texture index of image `image_id` (1-based, 0: none) under a sampler (address u, address v,
    //! min filter, mag filter); 0 for no image
    uint32_t add_texture(uint32_t image_id, const std::array<uint32_t, 4> &sampler)
    {
        if(!image_id) return 0;
        const std::array<uint32_t, 5> key{image_id - 1, sampler[0], sampler[1], sampler[2], sampler[3]};
        for(size_t k = 0; k < texture_refs.size(); ++k)
            if(texture_refs[k] == key) return static_cast<uint32_t>(k + 1);
        texture_refs.push_back(key);
        return static_cast<uint32_t>(texture_refs.size());
    }
    void finish()
    {
        textures.clear();
        for(const auto &r: texture_refs)
        {
            const image_rgba8 &im = images[r[0]];
            textures.push_back(vkrt_texture{im.width, im.height, im.rgba.data(), r[1], r[2], r[3], r[4]});
        }
    }

    //! append one object's geometry (indices relative to its own vertices)
    void add_object(const std::string &name, const std::vector<float> &pos, const std::vector<float> &nrm,
                    const std::vector<float> &uv, const std::vector<float> &tan, const std::vector<uint32_t> &idx,
                    uint32_t material)
    {
        const uint32_t base = static_cast<uint32_t>(positions.size() / 3);
        positions.insert(positions.end(), pos.begin(), pos.end());
        normals.insert(normals.end(), nrm.begin(), nrm.end());
        if(uv.size() == pos.size() / 3 * 2) texcoords.insert(texcoords.end(), uv.begin(), uv.end());
        else texcoords.resize(positions.size() / 3 * 2, 0.f);
        if(tan.size() == pos.size()) tangents.insert(tangents.end(), tan.begin(), tan.end());
        else tangents.resize(positions.size(), 0.f);
        for(uint32_t i: idx) indices.push_back(base + i);
        tri_object.resize(indices.size() / 3, n_objects);
        object_material.push_back(material);
        object_names.push_back(name);
        ++n_objects;
    }
};

namespace
{

// ---------------------------------------------------------------------------------------------
// Wavefront OBJ + MTL
// ---------------------------------------------------------------------------------------------
struct obj_material
{
    std::string name;
    float diffuse[3] = {0.f, 0.f, 0.f}, emission[3] = {0.f, 0.f, 0.f};
    float shininess = 1.f, ior = 1.f, dissolve = 1.f, roughness = 0.f, metallic = 0.f, clearcoat_roughness = 0.f;
    std::string diffuse_tex, normal_tex;
};

std::vector<obj_material> parse_mtl(const std::string &text)
{
    std::vector<obj_material> out;
    std::istringstream in(text);
    std::string line;
    while(std::getline(in, line))
    {
        if(!line.empty() && line.back() == '\r') line.pop_back();
        std::istringstream ls(line);
        std::string key;
        if(!(ls >> key) || key[0] == '#') continue;
        if(key == "newmtl")
        {
            out.emplace_back();
            std::getline(ls >> std::ws, out.back().name);
            continue;
        }
        if(out.empty()) continue;
        obj_material &m = out.back();
        auto f3 = [&ls](float *v) {
            for(int k = 0; k < 3; ++k) ls >> v[k];
        };
        if(key == "Kd") f3(m.diffuse);
        else if(key == "Ke") f3(m.emission);
        else if(key == "Ns") ls >> m.shininess;
        else if(key == "Ni") ls >> m.ior;
        else if(key == "d") ls >> m.dissolve;
        else if(key == "Tr")
        {
            float tr = 0.f;
            ls >> tr;
            m.dissolve = 1.f - tr;
        }
        else if(key == "Pr") ls >> m.roughness;
        else if(key == "Pm") ls >> m.metallic;
        else if(key == "Pcr") ls >> m.clearcoat_roughness;
        else if(key == "map_Kd") std::getline(ls >> std::ws, m.diffuse_tex);
        else if(key == "norm") std::getline(ls >> std::ws, m.normal_tex);// tinyobj's normal_texname
    }
    return out;
}

//! a face corner: 0-based indices, -1 when absent
struct corner
{
    int v = -1, vt = -1, vn = -1;
};

int resolve(long i, size_t count)
{
    if(i > 0) return static_cast<int>(i - 1);
    if(i < 0) return static_cast<int>(static_cast<long>(count) + i);
    return -1;
}

std::unique_ptr<vkrt_model> load_obj(const std::string &path)
{
    const std::string text = read_file(path), base = directory_of(path);
    std::vector<float> v, vt, vn;
    std::vector<obj_material> mtl;
    std::map<std::string, int> mtl_index;
    struct shape_t
    {
        std::string name;
        std::vector<corner> corners;// triangles, 3 per face
        int first_material = -2;     // -2: no face yet
    };
    std::vector<shape_t> shapes(1);
    int current_material = -1;
    std::istringstream in(text);
    std::string line;
    size_t line_no = 0;
    while(std::getline(in, line))
    {
        ++line_no;
        if(!line.empty() && line.back() == '\r') line.pop_back();
        std::istringstream ls(line);
        std::string key;
        if(!(ls >> key) || key[0] == '#') continue;
        if(key == "v")
        {
            float x = 0, y = 0, z = 0;
            if(!(ls >> x >> y >> z)) throw load_error(path + ":" + std::to_string(line_no) + ": bad vertex");
            v.insert(v.end(), {x, y, z});
        }
        else if(key == "vt")
        {
            float x = 0, y = 0;
            if(!(ls >> x)) throw load_error(path + ":" + std::to_string(line_no) + ": bad texture coordinate");
            ls >> y;
            vt.insert(vt.end(), {x, y});
        }
        else if(key == "vn")
        {
            float x = 0, y = 0, z = 0;
            if(!(ls >> x >> y >> z)) throw load_error(path + ":" + std::to_string(line_no) + ": bad normal");
            vn.insert(vn.end(), {x, y, z});
        }
        else if(key == "f")
        {
            std::vector<corner> poly;
            std::string tok;
            while(ls >> tok)
            {
                corner c;
                long a = 0, b = 0, d = 0;
                const size_t s1 = tok.find('/');
                a = std::strtol(tok.c_str(), nullptr, 10);
                if(s1 != std::string::npos)
                {
                    const size_t s2 = tok.find('/', s1 + 1);
                    if(s2 != s1 + 1) b = std::strtol(tok.c_str() + s1 + 1, nullptr, 10);
                    if(s2 != std::string::npos) d = std::strtol(tok.c_str() + s2 + 1, nullptr, 10);
                }
                c.v = resolve(a, v.size() / 3);
                c.vt = resolve(b, vt.size() / 2);
                c.vn = resolve(d, vn.size() / 3);
                if(c.v < 0 || c.v >= static_cast<int>(v.size() / 3) || c.vt >= static_cast<int>(vt.size() / 2) ||
                   c.vn >= static_cast<int>(vn.size() / 3))
                    throw load_error(path + ":" + std::to_string(line_no) + ": face index out of range");
                poly.push_back(c);
            }
            if(poly.size() < 3) continue;
            shape_t &sh = shapes.back();
            if(sh.first_material == -2) sh.first_material = current_material;
            for(size_t k = 1; k + 1 < poly.size(); ++k)// fan from the first corner
                sh.corners.insert(sh.corners.end(), {poly[0], poly[k], poly[k + 1]});
        }
        else if(key == "o" || key == "g")
        {
            std::string name;
            std::getline(ls >> std::ws, name);
            if(!shapes.back().corners.empty()) shapes.emplace_back();
            shapes.back().name = name;
        }
        else if(key == "usemtl")
        {
            std::string name;
            std::getline(ls >> std::ws, name);
            auto it = mtl_index.find(name);
            current_material = it == mtl_index.end() ? -1 : it->second;
        }
        else if(key == "mtllib")
        {
            std::string name;
            std::getline(ls >> std::ws, name);
            for(auto &m: parse_mtl(read_file(base + name)))
            {
                mtl_index[m.name] = static_cast<int>(mtl.size());
                mtl.push_back(m);
            }
        }
    }
    auto model = std::make_unique<vkrt_model>();
    std::map<std::string, uint32_t> tex_cache;
    // a material with a diffuse texture gets a texture transform that flips v (wavefront_obj.cpp:100-111); it
    // is applied here, to the coordinates of the faces that use it
    std::vector<bool> flip_v;
    auto load_texture = [&](const std::string &name) -> uint32_t {
        auto it = tex_cache.find(name);
        if(it == tex_cache.end())
        {
            uint32_t id = 0;
            std::ifstream f(base + name, std::ios::binary);
            if(f)
            {
                const std::string bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
                id = model->add_image(reinterpret_cast<const uint8_t *>(bytes.data()), bytes.size());
            }
            else ++model->n_skipped_textures;
            it = tex_cache.emplace(name, model->add_texture(id, {0, 0, 0, 0})).first;
        }
        return it->second;
    };
    for(const obj_material &m: mtl)
    {
        vkrt_material r = default_material();
        flip_v.push_back(!m.diffuse_tex.empty());
        if(!m.diffuse_tex.empty()) r.color_texture = load_texture(m.diffuse_tex);
        if(!m.normal_tex.empty()) r.normal_texture = load_texture(m.normal_tex);
        r.color[0] = m.diffuse[0];
        r.color[1] = m.diffuse[1];
        r.color[2] = m.diffuse[2];
        r.color[3] = std::fmin(std::fmax(1.f - m.dissolve, 0.f), 1.f);
        r.emission[0] = m.emission[0];
        r.emission[1] = m.emission[1];
        r.emission[2] = m.emission[2];
        r.roughness = std::fmin(std::fmax(std::fmax(m.roughness, std::pow(1.f - m.shininess, 2.f)), 0.f), 1.f);
        r.metalness = m.metallic;
        r.clearcoat_roughness = m.clearcoat_roughness;
        r.ior = m.ior;
        model->materials.push_back(r);
        model->material_names.push_back(m.name);
    }
    if(model->materials.empty())
    {
        model->materials.push_back(default_material());
        model->material_names.emplace_back();
    }
    const bool has_normals = !vn.empty();
    for(const shape_t &sh: shapes)
    {
        if(sh.corners.empty()) continue;
        const size_t n = sh.corners.size();
        const uint32_t material = static_cast<uint32_t>(std::max(sh.first_material, 0));
        const bool flip = material < flip_v.size() && flip_v[material];
        std::vector<float> pos(3 * n, 0.f), nrm, uv(2 * n, 0.f);
        std::vector<uint32_t> idx(n);
        if(has_normals) nrm.assign(3 * n, 0.f);
        for(size_t i = 0; i < n; ++i)
        {
            const corner &c = sh.corners[i];
            for(int a = 0; a < 3; ++a) pos[3 * i + a] = v[3 * (size_t)c.v + a];
            if(has_normals && c.vn >= 0)
                for(int a = 0; a < 3; ++a) nrm[3 * i + a] = vn[3 * (size_t)c.vn + a];
            if(c.vt >= 0)
            {
                uv[2 * i] = vt[2 * (size_t)c.vt];
                uv[2 * i + 1] = flip ? -vt[2 * (size_t)c.vt + 1] : vt[2 * (size_t)c.vt + 1];
            }
            idx[i] = static_cast<uint32_t>(i);
        }
        if(!has_normals) compute_vertex_normals(pos, idx, nrm);
        // texture coordinates are zero-filled per corner, so the reference computes tangents for every shape
        // (wavefront_obj.cpp:21, :44)
        std::vector<float> tan;
        compute_tangents(pos, nrm, uv, idx, tan);
        model->add_object(sh.name, pos, nrm, uv, tan, idx, material);
    }
    return model;
}

// ---------------------------------------------------------------------------------------------
// glTF 2.0
// ---------------------------------------------------------------------------------------------
std::vector<uint8_t> base64_decode(const std::string &s)
{
    std::vector<uint8_t> out;
    uint32_t acc = 0;
    int bits = 0;
    for(char c: s)
    {
        int d;
        if(c >= 'A' && c <= 'Z') d = c - 'A';
        else if(c >= 'a' && c <= 'z') d = c - 'a' + 26;
        else if(c >= '0' && c <= '9') d = c - '0' + 52;
        else if(c == '+' || c == '-') d = 62;
        else if(c == '/' || c == '_') d = 63;
        else if(c == '=') break;
        else continue;
        acc = (acc << 6) | static_cast<uint32_t>(d);
        bits += 6;
        if(bits >= 8)
        {
            bits -= 8;
            out.push_back(static_cast<uint8_t>((acc >> bits) & 0xFFu));
        }
    }
    return out;
}

struct gltf_file
{
    json doc;
    std::vector<std::vector<uint8_t>> buffers;
};

//! a JSON size field: absent gives `def`; negative, fractional, non-finite or above 2^40 values are rejected
size_t json_size(const json &v, size_t def, const char *what)
{
    if(v.kind != json::number_t) return def;
    const double x = v.num;
    if(!(x >= 0.0) || x > 1099511627776.0 || x != std::floor(x))
        throw load_error(std::string("glTF: invalid ") + what);
    return static_cast<size_t>(x);
}

gltf_file read_gltf(const std::string &path)
{
    const std::string bytes = read_file(path), base = directory_of(path);
    gltf_file g;
    std::vector<uint8_t> bin;
    const bool glb = bytes.size() >= 12 && std::memcmp(bytes.data(), "glTF", 4) == 0;
    if(glb)
    {
        uint32_t version, length;
        std::memcpy(&version, bytes.data() + 4, 4);
        std::memcpy(&length, bytes.data() + 8, 4);
        if(version != 2 || length > bytes.size()) throw load_error("GLB: bad header");
        size_t off = 12;
        bool have_json = false;
        while(off + 8 <= length)
        {
            uint32_t clen, ctype;
            std::memcpy(&clen, bytes.data() + off, 4);
            std::memcpy(&ctype, bytes.data() + off + 4, 4);
            if(off + 8 + clen > length) throw load_error("GLB: chunk past the end");
            const char *c = bytes.data() + off + 8;
            if(ctype == 0x4E4F534Au)
            {
                g.doc = parse_json(c, clen);
                have_json = true;
            }
            else if(ctype == 0x004E4942u) bin.assign(c, c + clen);
            off += 8 + ((clen + 3u) & ~3u);
        }
        if(!have_json) throw load_error("GLB: no JSON chunk");
    }
    else g.doc = parse_json(bytes.data(), bytes.size());
    const json &bufs = g.doc["buffers"];
    for(size_t i = 0; i < bufs.size(); ++i)
    {
        const json &b = bufs[i];
        const size_t len = json_size(b["byteLength"], 0, "buffer byteLength");
        std::vector<uint8_t> data;
        if(!b.has("uri"))
        {
            if(!glb || i != 0) throw load_error("glTF: buffer without uri");
            data = bin;
        }
        else
        {
            const std::string &uri = b["uri"].str;
            if(uri.rfind("data:", 0) == 0)
            {
                const size_t comma = uri.find(',');
                if(comma == std::string::npos || uri.find(";base64") == std::string::npos)
                    throw load_error("glTF: only base64 data URIs are supported");
                data = base64_decode(uri.substr(comma + 1));
            }
            else
            {
                const std::string s = read_file(base + uri);
                data.assign(s.begin(), s.end());
            }
        }
        if(data.size() < len) throw load_error("glTF: buffer shorter than its byteLength");
        g.buffers.push_back(std::move(data));
    }
    return g;
}

//! [offset, offset + length) of a bufferView inside its buffer, checked without overflow
struct view_range
{
    const std::vector<uint8_t> *buf;
    size_t offset, length;
};

view_range buffer_view(const gltf_file &g, const json &bv)
{
    if(bv.kind != json::object_t) throw load_error("glTF: bufferView out of range");
    const int bi = bv["buffer"].integer(-1);
    if(bi < 0 || bi >= static_cast<int>(g.buffers.size())) throw load_error("glTF: buffer out of range");
    const std::vector<uint8_t> &buf = g.buffers[bi];
    const size_t off = json_size(bv["byteOffset"], 0, "bufferView byteOffset");
    const size_t len = json_size(bv["byteLength"], 0, "bufferView byteLength");
    if(off > buf.size() || len > buf.size() - off) throw load_error("glTF: bufferView past its buffer");
    return {&buf, off, len};
}

//! one accessor as doubles (components per element in `comps`); float, and u8/u16/u32 for indices
std::vector<double> read_accessor(const gltf_file &g, int index, int &comps)
{
    const json &acc = g.doc["accessors"][index];
    if(acc.kind != json::object_t) throw load_error("glTF: accessor out of range");
    const std::string &type = acc["type"].str;
    comps = type == "SCALAR" ? 1 : type == "VEC2" ? 2 : type == "VEC3" ? 3 : type == "VEC4" ? 4 : 0;
    if(!comps) throw load_error("glTF: unsupported accessor type " + type);
    const int ct = acc["componentType"].integer(0);
    const size_t csize = ct == 5126 || ct == 5125 ? 4 : ct == 5123 || ct == 5122 ? 2 : ct == 5121 || ct == 5120 ? 1 : 0;
    if(!csize) throw load_error("glTF: unsupported component type");
    const size_t count = json_size(acc["count"], 0, "accessor count");
    if(count > (size_t(1) << 28)) throw load_error("glTF: accessor count above 2^28");
    if(!acc.has("bufferView")) return std::vector<double>(count * comps, 0.0);// sparse accessors are not supported
    const view_range view = buffer_view(g, g.doc["bufferViews"][acc["bufferView"].integer(-1)]);
    const size_t elem = csize * comps;
    const size_t stride = json_size(g.doc["bufferViews"][acc["bufferView"].integer(-1)]["byteStride"], elem,
                                    "bufferView byteStride");
    if(stride < elem) throw load_error("glTF: byteStride below the element size");
    const size_t acc_off = json_size(acc["byteOffset"], 0, "accessor byteOffset");
    // the last element ends at acc_off + (count - 1) stride + elem, which must lie inside the view; checked
    // as a division, so that no product can overflow
    if(count && (acc_off > view.length || elem > view.length - acc_off ||
                 (count - 1) > (view.length - acc_off - elem) / stride))
        throw load_error("glTF: accessor past its bufferView");
    std::vector<double> out(count * comps, 0.0);
    const uint8_t *start = view.buf->data() + view.offset + acc_off;
    for(size_t i = 0; i < count; ++i)
        for(int c = 0; c < comps; ++c)
        {
            const uint8_t *q = start + i * stride + c * csize;
            double x = 0;
            switch(ct)
            {
                case 5126: { float f; std::memcpy(&f, q, 4); x = f; break; }
                case 5125: { uint32_t u; std::memcpy(&u, q, 4); x = u; break; }
                case 5123: { uint16_t u; std::memcpy(&u, q, 2); x = u; break; }
                case 5122: { int16_t u; std::memcpy(&u, q, 2); x = u; break; }
                case 5121: x = *q; break;
                default: x = static_cast<int8_t>(*q); break;
            }
            out[i * comps + c] = x;
        }
    return out;
}

using mat4d = std::array<double, 16>;// column-major, as glTF stores it

mat4d identity4()
{
    mat4d m{};
    m[0] = m[5] = m[10] = m[15] = 1.0;
    return m;
}

mat4d mul4(const mat4d &a, const mat4d &b)
{
    mat4d r{};
    for(int c = 0; c < 4; ++c)
        for(int rr = 0; rr < 4; ++rr)
        {
            double s = 0;
            for(int k = 0; k < 4; ++k) s += a[k * 4 + rr] * b[c * 4 + k];
            r[c * 4 + rr] = s;
        }
    return r;
}

//! node_transform (gltf.cpp:136-160): the matrix, or translation * rotation * scale
mat4d node_matrix(const json &node)
{
    if(node["matrix"].size() == 16)
    {
        mat4d m{};
        for(int k = 0; k < 16; ++k) m[k] = node["matrix"][k].number(0);
        return m;
    }
    double t[3] = {0, 0, 0}, s[3] = {1, 1, 1}, q[4] = {0, 0, 0, 1};
    if(node["translation"].size() == 3)
        for(int k = 0; k < 3; ++k) t[k] = node["translation"][k].number(0);
    if(node["scale"].size() == 3)
        for(int k = 0; k < 3; ++k) s[k] = node["scale"][k].number(1);
    if(node["rotation"].size() == 4)
        for(int k = 0; k < 4; ++k) q[k] = node["rotation"][k].number(0);
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    const double r[9] = {1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w),
                         2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w),
                         2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)};// columns
    mat4d m{};
    for(int c = 0; c < 3; ++c)
        for(int rr = 0; rr < 3; ++rr) m[c * 4 + rr] = r[c * 3 + rr] * s[c];
    m[12] = t[0];
    m[13] = t[1];
    m[14] = t[2];
    m[15] = 1;
    return m;
}

//! KHR_texture_transform of the base color texture (texture_transform, gltf.cpp:162-194): uv' = offset +
//! R * (scale * uv) with R the rotation about -z, as (a, b, c, d, offset x, offset y) for
//! u' = a u + b v + ox, v' = c u + d v + oy. The path tracer applies it to every coordinate of the primitive
//! (the entry's texture_matrix, RayBuilder.cpp:596, raypipeline.slang:350); it is folded into the
//! coordinates at load time.
struct uv_transform
{
    bool active = false;
    float m[6] = {1.f, 0.f, 0.f, 1.f, 0.f, 0.f};
};

uv_transform base_color_transform(const json &jm)
{
    uv_transform r;
    const json &t = jm["pbrMetallicRoughness"]["baseColorTexture"]["extensions"]["KHR_texture_transform"];
    if(t.kind != json::object_t) return r;
    r.active = true;
    const float ox = static_cast<float>(t["offset"][0].number(0.0)), oy = static_cast<float>(t["offset"][1].number(0.0));
    const float rot = static_cast<float>(t["rotation"].number(0.0));
    const float sx = static_cast<float>(t["scale"][0].number(1.0)), sy = static_cast<float>(t["scale"][1].number(1.0));
    const float c = std::cos(rot), sn = std::sin(rot);
    r.m[0] = c * sx;
    r.m[1] = sn * sy;
    r.m[2] = -sn * sx;
    r.m[3] = c * sy;
    r.m[4] = ox;
    r.m[5] = oy;
    return r;
}

template<class F>
vkrt_material convert_material(const json &jm, F &&texture_of)
{
    vkrt_material m = default_material();
    const std::string &mode = jm["alphaMode"].str;
    m.alpha_mode = mode == "MASK" ? VKRT_ALPHA_MASK : mode == "BLEND" ? VKRT_ALPHA_BLEND : VKRT_ALPHA_OPAQUE;
    m.alpha_cutoff = static_cast<float>(jm["alphaCutoff"].number(0.5));
    if(jm["pbrMetallicRoughness"]["baseColorTexture"].has("index"))
        m.color_texture = texture_of(jm["pbrMetallicRoughness"]["baseColorTexture"]["index"].integer(-1));
    const json &pbr = jm["pbrMetallicRoughness"];
    for(int k = 0; k < 4; ++k) m.color[k] = static_cast<float>(pbr["baseColorFactor"][k].number(1.0));
    for(int k = 0; k < 3; ++k) m.emission[k] = static_cast<float>(jm["emissiveFactor"][k].number(0.0));
    m.metalness = static_cast<float>(pbr["metallicFactor"].number(1.0));
    m.roughness = static_cast<float>(pbr["roughnessFactor"].number(1.0));
    m.two_sided = jm["doubleSided"].kind == json::bool_t && jm["doubleSided"].b ? 1u : 0u;
    if(jm["emissiveTexture"].has("index"))
        for(int k = 0; k < 3; ++k) m.emission[k] = 0.f;// replaced by the texture (gltf.cpp:376)
    const json &ext = jm["extensions"];
    if(ext.has("KHR_materials_emissive_strength"))
        m.emission[3] = static_cast<float>(ext["KHR_materials_emissive_strength"]["emissiveStrength"].number(1.0));
    if(ext.has("KHR_materials_transmission"))
        m.transmission = static_cast<float>(ext["KHR_materials_transmission"]["transmissionFactor"].number(0.0));
    if(ext.has("KHR_materials_ior")) m.ior = static_cast<float>(ext["KHR_materials_ior"]["ior"].number(1.5));
    if(ext.has("KHR_materials_clearcoat"))
    {
        const json &c = ext["KHR_materials_clearcoat"];
        m.clearcoat = static_cast<float>(c["clearcoatFactor"].number(0.0));
        m.clearcoat_roughness = static_cast<float>(c["clearcoatRoughnessFactor"].number(0.0));
    }
    if(ext.has("KHR_materials_sheen"))
    {
        const json &sh = ext["KHR_materials_sheen"];
        for(int k = 0; k < 3; ++k) m.sheen_color[k] = static_cast<float>(sh["sheenColorFactor"][k].number(0.0));
        m.sheen_roughness = static_cast<float>(sh["sheenRoughnessFactor"].number(0.0));
    }
    if(ext.has("KHR_materials_diffuse_transmission"))
    {
        const json &dt = ext["KHR_materials_diffuse_transmission"];
        m.diffuse_transmission = static_cast<float>(dt["diffuseTransmissionFactor"].number(0.0));
        for(int k = 0; k < 3; ++k)
            m.diffuse_transmission_color[k] = static_cast<float>(dt["diffuseTransmissionColorFactor"][k].number(1.0));
    }
    if(ext.has("KHR_materials_volume"))
    {
        // attenuationColor and attenuationDistance as gltf.cpp:476-496 reads them (thickness is not used by the
        // tracer); the default distance +inf, no interior medium, is kept as 0, and a colour channel of 0 is
        // raised to 1e-6 (an extinction of 13.8 per attenuation distance) so that -log stays finite
        const json &vol = ext["KHR_materials_volume"];
        const double d = vol["attenuationDistance"].number(0.0);
        if(d > 0.0 && d < 3.4e38)
        {
            m.attenuation_distance = static_cast<float>(d);
            for(int k = 0; k < 3; ++k)
                m.attenuation_color[k] = static_cast<float>(
                        std::min(1.0, std::max(1e-6, vol["attenuationColor"][k].number(1.0))));
        }
    }
    // the other texture slots (gltf.cpp:366-378, :470-476, :622-632); occlusion only fills the red channel
    // of the rough/metal image, which the tracer does not read
    auto slot = [&](const json &info) -> uint32_t { return info.has("index") ? texture_of(info["index"].integer(-1)) : 0u; };
    m.ao_rough_metal_texture = slot(pbr["metallicRoughnessTexture"]);
    m.normal_texture = slot(jm["normalTexture"]);
    m.emission_texture = slot(jm["emissiveTexture"]);
    m.transmission_texture = slot(ext["KHR_materials_transmission"]["transmissionTexture"]);
    m.diffuse_transmission_texture = slot(ext["KHR_materials_diffuse_transmission"]["diffuseTransmissionTexture"]);
    m.diffuse_transmission_color_texture =
            slot(ext["KHR_materials_diffuse_transmission"]["diffuseTransmissionColorTexture"]);
    return m;
}

std::unique_ptr<vkrt_model> load_gltf(const std::string &path)
{
    gltf_file g = read_gltf(path);
    const json &doc = g.doc;
    auto model = std::make_unique<vkrt_model>();
    // images decode once, on first use by a material; a texture is an image with the glTF texture's sampler
    // (convert_sampler, gltf.cpp:662-690: NEAREST only for the exact NEAREST filter, the wrap modes as named)
    std::map<int, uint32_t> image_texture;
    auto sampler_of = [&](int tex) -> std::array<uint32_t, 4> {
        const int si = doc["textures"][static_cast<size_t>(tex)]["sampler"].integer(-1);
        const json &sm = doc["samplers"][static_cast<size_t>(si < 0 ? 0 : si)];
        if(si < 0 || sm.kind != json::object_t) return {0, 0, 0, 0};
        auto wrap = [](int w) -> uint32_t {
            return w == 33071 ? VKRT_ADDRESS_CLAMP_TO_EDGE : w == 33648 ? VKRT_ADDRESS_MIRRORED_REPEAT
                                                                        : VKRT_ADDRESS_REPEAT;
        };
        auto filter = [](int f) -> uint32_t { return f == 9728 ? VKRT_FILTER_NEAREST : VKRT_FILTER_LINEAR; };
        return {wrap(sm["wrapS"].integer(10497)), wrap(sm["wrapT"].integer(10497)), filter(sm["minFilter"].integer(-1)),
                filter(sm["magFilter"].integer(-1))};
    };
    auto texture_of = [&](int tex) -> uint32_t {
        const int src = doc["textures"][static_cast<size_t>(tex < 0 ? 0 : tex)]["source"].integer(-1);
        if(tex < 0 || src < 0) return 0u;
        auto it = image_texture.find(src);
        if(it != image_texture.end()) return model->add_texture(it->second, sampler_of(tex));
        const json &im = doc["images"][static_cast<size_t>(src)];
        uint32_t id = 0;
        if(im.has("bufferView"))
        {
            // an image whose bufferView is malformed is skipped like an undecodable one
            try
            {
                const view_range view =
                        buffer_view(g, doc["bufferViews"][static_cast<size_t>(im["bufferView"].integer(0))]);
                id = model->add_image(view.buf->data() + view.offset, view.length);
            }
            catch(const load_error &)
            {
                ++model->n_skipped_textures;
            }
        }
        else if(im.has("uri"))
        {
            const std::string &uri = im["uri"].str;
            std::vector<uint8_t> bytes;
            if(uri.rfind("data:", 0) == 0)
            {
                const size_t comma = uri.find(',');
                if(comma != std::string::npos) bytes = base64_decode(uri.substr(comma + 1));
            }
            else
            {
                std::ifstream f(directory_of(path) + uri, std::ios::binary);
                if(f) bytes.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
            }
            if(!bytes.empty()) id = model->add_image(bytes.data(), bytes.size());
            else ++model->n_skipped_textures;
        }
        image_texture[src] = id;
        return model->add_texture(id, sampler_of(tex));
    };
    const json &mats = doc["materials"];
    std::vector<uv_transform> uv_xform;
    for(size_t i = 0; i < mats.size(); ++i)
    {
        model->materials.push_back(convert_material(mats[i], texture_of));
        model->material_names.push_back(mats[i]["name"].str);
        uv_xform.push_back(model->materials.back().color_texture ? base_color_transform(mats[i]) : uv_transform{});
    }
    const uint32_t default_index = static_cast<uint32_t>(model->materials.size());
    bool default_used = false;

    // the scene's root nodes (the default scene, else the first, else every node that is no child)
    std::vector<int> roots;
    const json &scenes = doc["scenes"];
    if(scenes.size())
    {
        const json &sc = scenes[static_cast<size_t>(doc["scene"].integer(0))];
        for(size_t i = 0; i < sc["nodes"].size(); ++i) roots.push_back(sc["nodes"][i].integer(-1));
    }
    else
    {
        std::vector<bool> is_child(doc["nodes"].size(), false);
        for(size_t i = 0; i < doc["nodes"].size(); ++i)
            for(size_t k = 0; k < doc["nodes"][i]["children"].size(); ++k)
            {
                const int c = doc["nodes"][i]["children"][k].integer(-1);
                if(c >= 0 && c < static_cast<int>(is_child.size())) is_child[c] = true;
            }
        for(size_t i = 0; i < is_child.size(); ++i)
            if(!is_child[i]) roots.push_back(static_cast<int>(i));
    }
    struct item
    {
        int node;
        mat4d parent;
        int depth;
    };
    std::vector<item> todo;
    for(auto it = roots.rbegin(); it != roots.rend(); ++it) todo.push_back({*it, identity4(), 0});
    while(!todo.empty())
    {
        item cur = todo.back();
        todo.pop_back();
        const json &node = doc["nodes"][static_cast<size_t>(cur.node)];
        if(node.kind != json::object_t) throw load_error("glTF: node out of range");
        if(cur.depth > 1024) throw load_error("glTF: node hierarchy too deep (a cycle?)");
        const mat4d world = mul4(cur.parent, node_matrix(node));
        if(node.has("mesh"))
        {
            const json &mesh = doc["meshes"][static_cast<size_t>(node["mesh"].integer(-1))];
            if(mesh.kind != json::object_t) throw load_error("glTF: mesh out of range");
            // normals: inverse transpose of the upper 3x3
            const double a = world[0], b = world[4], c = world[8], d = world[1], e = world[5], f = world[9],
                         gg = world[2], h = world[6], i9 = world[10];
            const double det = a * (e * i9 - f * h) - b * (d * i9 - f * gg) + c * (d * h - e * gg);
            const double inv_t[9] = {(e * i9 - f * h) / det, -(d * i9 - f * gg) / det, (d * h - e * gg) / det,
                                     -(b * i9 - c * h) / det, (a * i9 - c * gg) / det, -(a * h - b * gg) / det,
                                     (b * f - c * e) / det, -(a * f - c * d) / det, (a * e - b * d) / det};
            for(size_t pi = 0; pi < mesh["primitives"].size(); ++pi)
            {
                const json &prim = mesh["primitives"][pi];
                if(prim["mode"].integer(4) != 4) continue;// triangles only
                const json &attr = prim["attributes"];
                if(!attr.has("POSITION")) continue;
                int comps = 0;
                const std::vector<double> p = read_accessor(g, attr["POSITION"].integer(-1), comps);
                if(comps != 3) throw load_error("glTF: POSITION must be VEC3");
                const size_t nv = p.size() / 3;
                std::vector<uint32_t> idx;
                if(prim.has("indices"))
                {
                    const std::vector<double> ix = read_accessor(g, prim["indices"].integer(-1), comps);
                    for(double x: ix)
                    {
                        if(x >= static_cast<double>(nv)) throw load_error("glTF: index out of range");
                        idx.push_back(static_cast<uint32_t>(x));
                    }
                }
                else
                    for(size_t k = 0; k < nv; ++k) idx.push_back(static_cast<uint32_t>(k));
                idx.resize(idx.size() / 3 * 3);
                std::vector<float> pos(3 * nv), nrm, uv;
                for(size_t k = 0; k < nv; ++k)
                    for(int r = 0; r < 3; ++r)
                        pos[3 * k + r] = static_cast<float>(world[r] * p[3 * k] + world[4 + r] * p[3 * k + 1] +
                                                            world[8 + r] * p[3 * k + 2] + world[12 + r]);
                if(attr.has("NORMAL"))
                {
                    const std::vector<double> n = read_accessor(g, attr["NORMAL"].integer(-1), comps);
                    if(comps != 3 || n.size() != p.size()) throw load_error("glTF: NORMAL must match POSITION");
                    nrm.resize(3 * nv);
                    for(size_t k = 0; k < nv; ++k)
                    {
                        double w[3];
                        for(int r = 0; r < 3; ++r)
                            w[r] = inv_t[r] * n[3 * k] + inv_t[3 + r] * n[3 * k + 1] + inv_t[6 + r] * n[3 * k + 2];
                        // renormalized only where the transform scales them, so identity nodes keep the
                        // file's bits
                        const double l = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
                        const double sc = (l > 0 && std::fabs(l - 1.0) > 1e-6) ? 1.0 / l : 1.0;
                        for(int r = 0; r < 3; ++r) nrm[3 * k + r] = static_cast<float>(w[r] * sc);
                    }
                }
                else
                    compute_vertex_normals(pos, idx, nrm);
                if(attr.has("TEXCOORD_0"))
                {
                    const std::vector<double> t = read_accessor(g, attr["TEXCOORD_0"].integer(-1), comps);
                    if(comps == 2 && t.size() == 2 * nv)
                        for(double x: t) uv.push_back(static_cast<float>(x));
                }
                // tangents: the file's TANGENT (its xyz; create_geometry's strided copy drops w) turned by the
                // node, else Geometry::compute_tangents from the file's coordinates (gltf.cpp:277, :302)
                std::vector<float> tan;
                if(attr.has("TANGENT"))
                {
                    const std::vector<double> tv = read_accessor(g, attr["TANGENT"].integer(-1), comps);
                    if((comps == 3 || comps == 4) && tv.size() == nv * comps)
                    {
                        tan.resize(3 * nv);
                        for(size_t k = 0; k < nv; ++k)
                        {
                            const double *t = tv.data() + k * comps;
                            double w[3];
                            for(int r = 0; r < 3; ++r) w[r] = world[r] * t[0] + world[4 + r] * t[1] + world[8 + r] * t[2];
                            const double l = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
                            const double sc = (l > 0 && std::fabs(l - 1.0) > 1e-6) ? 1.0 / l : 1.0;
                            for(int r = 0; r < 3; ++r) tan[3 * k + r] = static_cast<float>(w[r] * sc);
                        }
                    }
                }
                if(tan.empty() && !uv.empty()) compute_tangents(pos, nrm, uv, idx, tan);
                int mi = prim["material"].integer(-1);
                if(mi < 0 || mi >= static_cast<int>(default_index))
                {
                    mi = static_cast<int>(default_index);
                    default_used = true;
                }
                else if(uv_xform[mi].active)
                {
                    const float *m = uv_xform[mi].m;
                    for(size_t k = 0; k + 1 < uv.size(); k += 2)
                    {
                        const float u = uv[k], v = uv[k + 1];
                        uv[k] = (m[0] * u + m[1] * v) + m[4];
                        uv[k + 1] = (m[2] * u + m[3] * v) + m[5];
                    }
                }
                model->add_object(mesh["name"].str, pos, nrm, uv, tan, idx, static_cast<uint32_t>(mi));
            }
        }
        const json &ch = node["children"];
        for(size_t k = ch.size(); k-- > 0;) todo.push_back({ch[k].integer(-1), world, cur.depth + 1});
    }
    if(default_used || model->materials.empty())
    {
        model->materials.push_back(default_material());
        model->material_names.emplace_back();
    }
    return model;
}

}// namespace

extern "C" {

int vkrt_model_load(const char *path, vkrt_model **out)
{
    if(!out) return VKRT_E_INVALID_ARG;
    *out = nullptr;
    try
    {
        if(!path) throw load_error("null path");
        const std::string p(path), ext = lower_extension(p);
        std::unique_ptr<vkrt_model> m;
        if(ext == "obj") m = load_obj(p);
        else if(ext == "gltf" || ext == "glb") m = load_gltf(p);
        else throw load_error("unknown model format: " + p);
        if(m->n_objects == 0)
        {
            // an empty scene still needs one object and one material
            m->object_material.push_back(0);
            m->object_names.emplace_back();
            m->n_objects = 1;
        }
        m->finish();
        g_model_error.clear();
        *out = m.release();
        return VKRT_OK;
    } catch(const load_error &e)
    {
        g_model_error = e.what();
        return VKRT_E_INVALID_ARG;
    } catch(const std::bad_alloc &)
    {
        g_model_error = "host allocation failed";
        return VKRT_E_OOM;
    } catch(const std::exception &e)
    {
        g_model_error = e.what();
        return VKRT_E_INTERNAL;
    }
}

const char *vkrt_model_error(void) { return g_model_error.c_str(); }

int vkrt_model_get(const vkrt_model *m, vkrt_model_view *out)
{
    if(!m || !out) return VKRT_E_INVALID_ARG;
    *out = {};
    out->positions = m->positions.data();
    out->normals = m->normals.data();
    out->texcoords = m->texcoords.data();
    out->tangents = m->tangents.data();
    out->n_vert = static_cast<uint32_t>(m->positions.size() / 3);
    out->indices = m->indices.data();
    out->n_tri = static_cast<uint32_t>(m->indices.size() / 3);
    out->tri_object = m->tri_object.data();
    out->object_material = m->object_material.data();
    out->n_objects = m->n_objects;
    out->materials = m->materials.data();
    out->n_materials = static_cast<uint32_t>(m->materials.size());
    out->textures = m->textures.data();
    out->n_textures = static_cast<uint32_t>(m->textures.size());
    out->n_skipped_textures = m->n_skipped_textures;
    return VKRT_OK;
}

int vkrt_model_pt_scene(const vkrt_model *m, float environment, vkrt_pt_scene *out)
{
    vkrt_model_view v;
    const int rc = vkrt_model_get(m, &v);
    if(rc != VKRT_OK) return rc;
    *out = vkrt_pt_scene{v.positions, v.normals, v.n_vert, v.indices, v.n_tri, v.tri_object, v.object_material,
                         v.n_objects, v.materials, v.n_materials, environment, v.texcoords, v.textures,
                         v.n_textures, v.tangents};
    return VKRT_OK;
}

void vkrt_model_destroy(vkrt_model *m) { delete m; }

}// extern "C"
