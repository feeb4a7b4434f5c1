// gltf.cpp - glTF 2.0 scene loading into the C ABI's host arrays (FSceneLoader::
// LoadGLTFScene, src/Assets/FSceneLoader.cpp:102-700). Host code only: a JSON reader, the
// GLB container, buffers from files or base64 data URIs, a PNG decoder over zlib, and the
// reference's material, node, light and camera rules.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/gknx_rt.h"

namespace {

thread_local std::string g_err;

struct Fail {
    int code;
};
[[noreturn]] void fail(int code, const std::string& msg) {
    g_err = msg;
    throw Fail{code};
}

// ---------------------------------------------------------------- JSON
struct JVal {
    enum Type { Null, Bool, Num, Str, Arr, Obj } t = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;

    const JVal* get(const char* key) const {
        if (t != Obj) return nullptr;
        for (const auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    double num_or(const char* key, double d) const {
        const JVal* v = get(key);
        return v && v->t == Num ? v->num : d;
    }
    int int_or(const char* key, int d) const { return int(num_or(key, d)); }
    size_t size() const { return t == Arr ? arr.size() : 0; }
    const JVal& at(size_t i) const {
        if (t != Arr || i >= arr.size()) fail(GK_ERR_ARG, "glTF: index out of range");
        return arr[i];
    }
};

struct JParser {
    const char* p;
    const char* e;
    void ws() {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    [[noreturn]] void bad() { fail(GK_ERR_ARG, "glTF: malformed JSON"); }
    void expect(char c) {
        ws();
        if (p >= e || *p != c) bad();
        ++p;
    }
    static void utf8(std::string& o, uint32_t cp) {
        if (cp < 0x80) o += char(cp);
        else if (cp < 0x800) { o += char(0xC0 | (cp >> 6)); o += char(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) {
            o += char(0xE0 | (cp >> 12)); o += char(0x80 | ((cp >> 6) & 0x3F)); o += char(0x80 | (cp & 0x3F));
        } else {
            o += char(0xF0 | (cp >> 18)); o += char(0x80 | ((cp >> 12) & 0x3F));
            o += char(0x80 | ((cp >> 6) & 0x3F)); o += char(0x80 | (cp & 0x3F));
        }
    }
    uint32_t hex4() {
        if (e - p < 4) bad();
        uint32_t v = 0;
        for (int i = 0; i < 4; ++i) {
            const char c = *p++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= uint32_t(c - '0');
            else if (c >= 'a' && c <= 'f') v |= uint32_t(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= uint32_t(c - 'A' + 10);
            else bad();
        }
        return v;
    }
    std::string string() {
        expect('"');
        std::string o;
        while (true) {
            if (p >= e) bad();
            const char c = *p++;
            if (c == '"') break;
            if (c != '\\') { o += c; continue; }
            if (p >= e) bad();
            const char x = *p++;
            switch (x) {
                case '"': o += '"'; break;
                case '\\': o += '\\'; break;
                case '/': o += '/'; break;
                case 'b': o += '\b'; break;
                case 'f': o += '\f'; break;
                case 'n': o += '\n'; break;
                case 'r': o += '\r'; break;
                case 't': o += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00 && e - p >= 6 && p[0] == '\\' && p[1] == 'u') {
                        p += 2;
                        const uint32_t lo = hex4();
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    utf8(o, cp);
                    break;
                }
                default: bad();
            }
        }
        return o;
    }
    JVal value(int depth = 0) {
        if (depth > 256) bad();
        ws();
        if (p >= e) bad();
        JVal v;
        if (*p == '{') {
            ++p;
            v.t = JVal::Obj;
            ws();
            if (p < e && *p == '}') { ++p; return v; }
            while (true) {
                std::string k = string();
                expect(':');
                v.obj.emplace_back(std::move(k), value(depth + 1));
                ws();
                if (p < e && *p == ',') { ++p; continue; }
                expect('}');
                return v;
            }
        }
        if (*p == '[') {
            ++p;
            v.t = JVal::Arr;
            ws();
            if (p < e && *p == ']') { ++p; return v; }
            while (true) {
                v.arr.push_back(value(depth + 1));
                ws();
                if (p < e && *p == ',') { ++p; continue; }
                expect(']');
                return v;
            }
        }
        if (*p == '"') { v.t = JVal::Str; v.str = string(); return v; }
        if (e - p >= 4 && std::strncmp(p, "true", 4) == 0) { p += 4; v.t = JVal::Bool; v.b = true; return v; }
        if (e - p >= 5 && std::strncmp(p, "false", 5) == 0) { p += 5; v.t = JVal::Bool; return v; }
        if (e - p >= 4 && std::strncmp(p, "null", 4) == 0) { p += 4; return v; }
        const char* s = p;
        while (p < e && (std::strchr("+-0123456789.eE", *p) != nullptr)) ++p;
        if (p == s) bad();
        v.t = JVal::Num;
        v.num = std::strtod(std::string(s, p).c_str(), nullptr);
        return v;
    }
};

// ---------------------------------------------------------------- bytes
std::vector<uint8_t> read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail(GK_ERR_IO, "glTF: cannot open " + path);
    return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

std::vector<uint8_t> base64(const std::string& s) {
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+' || c == '-') return 62;
        if (c == '/' || c == '_') return 63;
        return -1;
    };
    std::vector<uint8_t> o;
    uint32_t acc = 0;
    int bits = 0;
    for (const char c : s) {
        const int v = val(c);
        if (v < 0) continue;  // padding, whitespace
        acc = (acc << 6) | uint32_t(v);
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            o.push_back(uint8_t((acc >> bits) & 0xFF));
        }
    }
    return o;
}

std::vector<uint8_t> uri_bytes(const std::string& uri, const std::string& dir) {
    if (uri.compare(0, 5, "data:") == 0) {
        const size_t comma = uri.find(',');
        if (comma == std::string::npos || uri.find(";base64") > comma) fail(GK_ERR_ARG, "glTF: unsupported data URI");
        return base64(uri.substr(comma + 1));
    }
    std::string path;
    for (size_t i = 0; i < uri.size(); ++i) {  // percent-decoding of file names
        if (uri[i] == '%' && i + 2 < uri.size()) {
            path += char(std::strtol(uri.substr(i + 1, 2).c_str(), nullptr, 16));
            i += 2;
        } else {
            path += uri[i];
        }
    }
    return read_file(dir.empty() ? path : dir + "/" + path);
}

// ---------------------------------------------------------------- PNG
uint32_t be32(const uint8_t* b) { return (uint32_t(b[0]) << 24) | (uint32_t(b[1]) << 16) | (uint32_t(b[2]) << 8) | b[3]; }

// PNG (ISO/IEC 15948): 8- and 16-bit grey, RGB, palette, grey+alpha and RGBA, not
// interlaced, to RGBA8 (16-bit samples keep their high byte).
void decode_png(const std::vector<uint8_t>& d, uint32_t& w, uint32_t& h, std::vector<uint8_t>& rgba) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (d.size() < 8 || std::memcmp(d.data(), sig, 8) != 0) fail(GK_ERR_ARG, "glTF: image is not PNG (JPEG is not supported)");
    size_t p = 8;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    while (p + 12 <= d.size()) {
        const uint32_t len = be32(&d[p]);
        if (p + 12 + size_t(len) > d.size()) fail(GK_ERR_ARG, "glTF: truncated PNG");
        const char* type = reinterpret_cast<const char*>(&d[p + 4]);
        const uint8_t* body = &d[p + 8];
        if (std::strncmp(type, "IHDR", 4) == 0 && len >= 13) {
            w = be32(body);
            h = be32(body + 4);
            depth = body[8];
            ctype = body[9];
            interlace = body[12];
        } else if (std::strncmp(type, "PLTE", 4) == 0) {
            plte.assign(body, body + len);
        } else if (std::strncmp(type, "tRNS", 4) == 0) {
            trns.assign(body, body + len);
        } else if (std::strncmp(type, "IDAT", 4) == 0) {
            idat.insert(idat.end(), body, body + len);
        } else if (std::strncmp(type, "IEND", 4) == 0) {
            break;
        }
        p += 12 + size_t(len);
    }
    if (w == 0 || h == 0 || w > 16384 || h > 16384) fail(GK_ERR_ARG, "glTF: PNG size");
    if (interlace != 0) fail(GK_ERR_ARG, "glTF: interlaced PNG is not supported");
    int chans = 0;
    switch (ctype) {
        case 0: chans = 1; break;
        case 2: chans = 3; break;
        case 3: chans = 1; break;
        case 4: chans = 2; break;
        case 6: chans = 4; break;
        default: fail(GK_ERR_ARG, "glTF: PNG colour type");
    }
    if (!(depth == 8 || depth == 16 || (ctype == 3 && (depth == 1 || depth == 2 || depth == 4))))
        fail(GK_ERR_ARG, "glTF: PNG bit depth");
    const size_t bits_px = size_t(chans) * depth;
    const size_t stride = (size_t(w) * bits_px + 7) / 8;
    const size_t bpp = std::max<size_t>(1, bits_px / 8);
    std::vector<uint8_t> raw((stride + 1) * h);
    z_stream zs{};
    if (inflateInit(&zs) != Z_OK) fail(GK_ERR_ARG, "glTF: zlib");
    zs.next_in = idat.data();
    zs.avail_in = uInt(idat.size());
    zs.next_out = raw.data();
    zs.avail_out = uInt(raw.size());
    const int zr = inflate(&zs, Z_FINISH);
    const size_t got = raw.size() - zs.avail_out;
    inflateEnd(&zs);
    if ((zr != Z_STREAM_END && zr != Z_OK && zr != Z_BUF_ERROR) || got != raw.size())
        fail(GK_ERR_ARG, "glTF: corrupt PNG data");
    std::vector<uint8_t> img(stride * h), prev(stride, 0);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t f = raw[y * (stride + 1)];
        const uint8_t* in = &raw[y * (stride + 1) + 1];
        uint8_t* out = &img[y * stride];
        for (size_t x = 0; x < stride; ++x) {
            const int a = x >= bpp ? out[x - bpp] : 0, b = prev[x], c = x >= bpp ? prev[x - bpp] : 0;
            int v = in[x];
            switch (f) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) >> 1; break;
                case 4: {
                    const int pp = a + b - c, pa = std::abs(pp - a), pb = std::abs(pp - b), pc = std::abs(pp - c);
                    v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                    break;
                }
                default: fail(GK_ERR_ARG, "glTF: PNG filter");
            }
            out[x] = uint8_t(v);
        }
        std::memcpy(prev.data(), out, stride);
    }
    rgba.assign(size_t(w) * h * 4, 255);
    auto sample = [&](const uint8_t* row, size_t x, int ch) -> uint32_t {
        if (depth == 8) return row[x * chans + ch];
        if (depth == 16) return row[(x * chans + ch) * 2];
        const size_t bit = x * depth;  // palette indices of 1, 2 or 4 bits
        return (row[bit / 8] >> (8 - depth - int(bit % 8))) & ((1u << depth) - 1u);
    };
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t* row = &img[y * stride];
        for (uint32_t x = 0; x < w; ++x) {
            uint8_t* o = &rgba[(size_t(y) * w + x) * 4];
            if (ctype == 3) {
                const uint32_t i = sample(row, x, 0);
                if (3 * size_t(i) + 2 >= plte.size()) fail(GK_ERR_ARG, "glTF: PNG palette index");
                o[0] = plte[3 * i]; o[1] = plte[3 * i + 1]; o[2] = plte[3 * i + 2];
                o[3] = i < trns.size() ? trns[i] : 255;
            } else if (ctype == 0 || ctype == 4) {
                o[0] = o[1] = o[2] = uint8_t(sample(row, x, 0));
                if (ctype == 4) o[3] = uint8_t(sample(row, x, 1));
            } else {
                for (int k = 0; k < 3; ++k) o[k] = uint8_t(sample(row, x, k));
                if (ctype == 6) o[3] = uint8_t(sample(row, x, 3));
            }
        }
    }
}

// ---------------------------------------------------------------- scene
struct Mat4 {
    float m[16];  // column-major
};
Mat4 mul(const Mat4& a, const Mat4& b) {
    Mat4 o;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            float s = 0.f;
            for (int k = 0; k < 4; ++k) s += a.m[k * 4 + r] * b.m[c * 4 + k];
            o.m[c * 4 + r] = s;
        }
    return o;
}
void xform(const Mat4& M, const float v[4], float o[4]) {
    for (int r = 0; r < 4; ++r)
        o[r] = M.m[r] * v[0] + M.m[4 + r] * v[1] + M.m[8 + r] * v[2] + M.m[12 + r] * v[3];
}

struct Scene {
    JVal doc;
    std::string dir;
    std::vector<std::vector<uint8_t>> buffers;
    // results
    std::vector<std::vector<float>> pos, nrm, uv, tan;
    std::vector<std::vector<uint32_t>> idx, tmat;
    std::vector<gk_mesh> meshes;
    std::vector<gk_material> materials;
    std::vector<gk_instance> instances;
    std::vector<gk_light> lights;
    std::vector<std::vector<uint8_t>> tex_data;
    std::vector<gk_texture> textures;
    gk_gltf_view view{};

    const JVal& arr(const char* key) const {
        static const JVal empty;
        const JVal* v = doc.get(key);
        return v ? *v : empty;
    }

    // one accessor as floats (normalized integers converted as glTF specifies)
    std::vector<float> floats(int acc, int comps) {
        const JVal& a = arr("accessors").at(size_t(acc));
        if (a.get("sparse")) fail(GK_ERR_ARG, "glTF: sparse accessors are not supported");
        const size_t count = size_t(a.num_or("count", 0));
        const int ctype = a.int_or("componentType", 5126);
        const bool normalized = a.get("normalized") && a.get("normalized")->b;
        int csize = ctype == 5126 || ctype == 5125 ? 4 : (ctype == 5123 || ctype == 5122 ? 2 : 1);
        std::vector<float> out(count * comps, 0.f);
        const JVal* bv_idx = a.get("bufferView");
        if (!bv_idx) return out;  // no data: zeros
        const JVal& bv = arr("bufferViews").at(size_t(bv_idx->num));
        const size_t stride = size_t(bv.num_or("byteStride", 0)) ? size_t(bv.num_or("byteStride", 0)) : size_t(csize) * comps;
        const size_t off = size_t(bv.num_or("byteOffset", 0)) + size_t(a.num_or("byteOffset", 0));
        const size_t buf = size_t(bv.num_or("buffer", 0));
        if (buf >= buffers.size()) fail(GK_ERR_ARG, "glTF: buffer index");
        const std::vector<uint8_t>& B = buffers[buf];
        if (count && off + (count - 1) * stride + size_t(csize) * comps > B.size()) fail(GK_ERR_ARG, "glTF: accessor out of range");
        for (size_t i = 0; i < count; ++i)
            for (int k = 0; k < comps; ++k) {
                const uint8_t* q = &B[off + i * stride + size_t(k) * csize];
                float v;
                switch (ctype) {
                    case 5126: std::memcpy(&v, q, 4); break;
                    case 5121: v = normalized ? q[0] / 255.f : float(q[0]); break;
                    case 5120: v = normalized ? std::max(int8_t(q[0]) / 127.f, -1.f) : float(int8_t(q[0])); break;
                    case 5123: { uint16_t u; std::memcpy(&u, q, 2); v = normalized ? u / 65535.f : float(u); break; }
                    case 5122: { int16_t u; std::memcpy(&u, q, 2); v = normalized ? std::max(u / 32767.f, -1.f) : float(u); break; }
                    case 5125: { uint32_t u; std::memcpy(&u, q, 4); v = float(u); break; }
                    default: fail(GK_ERR_ARG, "glTF: component type");
                }
                out[i * comps + k] = v;
            }
        return out;
    }
    std::vector<uint32_t> indices(int acc) {
        const JVal& a = arr("accessors").at(size_t(acc));
        const size_t count = size_t(a.num_or("count", 0));
        const int ctype = a.int_or("componentType", 5125);
        const JVal* bv_idx = a.get("bufferView");
        std::vector<uint32_t> out(count, 0);
        if (!bv_idx) return out;
        const JVal& bv = arr("bufferViews").at(size_t(bv_idx->num));
        const size_t csize = ctype == 5121 ? 1 : (ctype == 5123 ? 2 : 4);
        const size_t stride = size_t(bv.num_or("byteStride", 0)) ? size_t(bv.num_or("byteStride", 0)) : csize;
        const size_t off = size_t(bv.num_or("byteOffset", 0)) + size_t(a.num_or("byteOffset", 0));
        const size_t buf = size_t(bv.num_or("buffer", 0));
        if (buf >= buffers.size()) fail(GK_ERR_ARG, "glTF: buffer index");
        const std::vector<uint8_t>& B = buffers[buf];
        if (count && off + (count - 1) * stride + csize > B.size()) fail(GK_ERR_ARG, "glTF: index accessor out of range");
        for (size_t i = 0; i < count; ++i) {
            const uint8_t* q = &B[off + i * stride];
            if (ctype == 5121) out[i] = q[0];
            else if (ctype == 5123) { uint16_t u; std::memcpy(&u, q, 2); out[i] = u; }
            else if (ctype == 5125 || ctype == 5124) { uint32_t u; std::memcpy(&u, q, 4); out[i] = u; }
            else fail(GK_ERR_ARG, "glTF: index component type");
        }
        return out;
    }

    std::vector<uint8_t> image_bytes(const JVal& img) {
        if (const JVal* bv = img.get("bufferView")) {
            const JVal& v = arr("bufferViews").at(size_t(bv->num));
            const size_t buf = size_t(v.num_or("buffer", 0)), off = size_t(v.num_or("byteOffset", 0));
            const size_t len = size_t(v.num_or("byteLength", 0));
            if (buf >= buffers.size() || off + len > buffers[buf].size()) fail(GK_ERR_ARG, "glTF: image buffer view");
            return std::vector<uint8_t>(buffers[buf].begin() + off, buffers[buf].begin() + off + len);
        }
        const JVal* uri = img.get("uri");
        if (!uri) fail(GK_ERR_ARG, "glTF: image without data");
        return uri_bytes(uri->str, dir);
    }

    void build() {
        // textures in the reference's load order: per material base colour (sRGB),
        // metallic-roughness, normal (FSceneLoader.cpp:304-309)
        const JVal& mats = arr("materials");
        std::vector<int> tex_of_image(arr("images").size(), -1);
        auto image_of = [&](int texture) {
            const JVal& t = arr("textures").at(size_t(texture));
            int img = t.int_or("source", -1);
            if (img == -1) img = texture;
            if (img < 0 || size_t(img) >= tex_of_image.size()) fail(GK_ERR_ARG, "glTF: texture source");
            return img;
        };
        auto load_tex = [&](const JVal* info, bool srgb) {
            if (!info) return;
            const int texture = info->int_or("index", -1);
            if (texture < 0) return;
            const int img = image_of(texture);
            if (tex_of_image[img] != -1) return;
            gk_texture t{};
            std::vector<uint8_t> rgba;
            decode_png(image_bytes(arr("images").at(size_t(img))), t.width, t.height, rgba);
            t.srgb = srgb ? 1u : 0u;
            tex_data.push_back(std::move(rgba));
            textures.push_back(t);
            tex_of_image[img] = int(textures.size() - 1);
        };
        for (size_t i = 0; i < mats.size(); ++i) {
            const JVal& m = mats.at(i);
            const JVal* pbr = m.get("pbrMetallicRoughness");
            load_tex(pbr ? pbr->get("baseColorTexture") : nullptr, true);
            load_tex(pbr ? pbr->get("metallicRoughnessTexture") : nullptr, false);
            load_tex(m.get("normalTexture"), false);
        }
        for (size_t i = 0; i < textures.size(); ++i) textures[i].rgba8 = tex_data[i].data();
        auto tex_id = [&](const JVal* info) {
            if (!info) return -1;
            const int texture = info->int_or("index", -1);
            return texture < 0 ? -1 : tex_of_image[image_of(texture)];
        };
        // materials (FSceneLoader.cpp:312-388)
        for (size_t i = 0; i < mats.size(); ++i) {
            const JVal& src = mats.at(i);
            const JVal* pbr = src.get("pbrMetallicRoughness");
            gk_material m{};
            m.diffuse_texture_id = m.mra_texture_id = m.normal_texture_id = -1;
            m.normal_texture_scale = 1.0f;
            m.material_model = GK_MATERIAL_MIXTURE;
            m.fuzziness = pbr ? float(pbr->num_or("roughnessFactor", 1.0)) : 1.0f;
            m.metalness = pbr ? float(pbr->num_or("metallicFactor", 1.0)) : 1.0f;
            m.refraction_index = m.refraction_index2 = 1.46f;
            m.diffuse_texture_id = tex_id(pbr ? pbr->get("baseColorTexture") : nullptr);
            m.mra_texture_id = tex_id(pbr ? pbr->get("metallicRoughnessTexture") : nullptr);
            m.normal_texture_id = tex_id(src.get("normalTexture"));
            if (const JVal* nt = src.get("normalTexture")) m.normal_texture_scale = float(nt->num_or("scale", 1.0));
            float emissive[3] = {0.f, 0.f, 0.f}, base[3] = {1.f, 1.f, 1.f};
            if (const JVal* ef = src.get("emissiveFactor"))
                for (size_t k = 0; k < 3 && k < ef->size(); ++k) emissive[k] = float(ef->at(k).num);
            if (const JVal* bc = pbr ? pbr->get("baseColorFactor") : nullptr)
                for (size_t k = 0; k < 3 && k < bc->size(); ++k) base[k] = float(bc->at(k).num);
            for (int k = 0; k < 3; ++k) m.diffuse[k] = std::sqrt(base[k]);
            m.diffuse[3] = 1.0f;
            if (m.mra_texture_id != -1) m.fuzziness = 1.0f;
            if (m.metalness > .95f && m.mra_texture_id == -1) m.material_model = GK_MATERIAL_METALLIC;
            if (m.fuzziness > .95f && m.mra_texture_id == -1 && m.metalness < 0.01f) m.material_model = GK_MATERIAL_LAMBERTIAN;
            const JVal* ext = src.get("extensions");
            if (const JVal* ior = ext ? ext->get("KHR_materials_ior") : nullptr) {
                m.refraction_index = float(ior->num_or("ior", 1.5));
                m.refraction_index2 = m.refraction_index;
            }
            if (const JVal* tr = ext ? ext->get("KHR_materials_transmission") : nullptr)
                if (float(tr->num_or("transmissionFactor", 0.0)) > 0.8f) m.material_model = GK_MATERIAL_DIELECTRIC;
            if (const JVal* ex = src.get("extras"))
                if (ex->get("ior2")) m.refraction_index2 = float(ex->num_or("ior2", 1.46));
            if (const JVal* es = ext ? ext->get("KHR_materials_emissive_strength") : nullptr) {
                const float power = float(es->num_or("emissiveStrength", 1.0));
                gk_material l{};  // Material::DiffuseLight (Material.hpp:35-38)
                for (int k = 0; k < 3; ++k) l.diffuse[k] = emissive[k] * power * 100.0f;
                l.diffuse[3] = 1.0f;
                l.diffuse_texture_id = -1;
                l.mra_texture_id = l.normal_texture_id = -1;
                l.material_model = GK_MATERIAL_DIFFUSE_LIGHT;
                m = l;
            }
            materials.push_back(m);
        }
        if (materials.empty()) {
            // no materials: primitives use slot max(0, -1) = 0; give it the mapping of an
            // empty glTF material (base colour 1, roughness 1, metallic 1 -> Metallic)
            gk_material m{};
            m.diffuse[0] = m.diffuse[1] = m.diffuse[2] = m.diffuse[3] = 1.f;
            m.diffuse_texture_id = m.mra_texture_id = m.normal_texture_id = -1;
            m.fuzziness = m.metalness = 1.f;
            m.refraction_index = m.refraction_index2 = 1.46f;
            m.normal_texture_scale = 1.f;
            m.material_model = GK_MATERIAL_METALLIC;
            materials.push_back(m);
        }
        // meshes (FSceneLoader.cpp:390-510)
        const JVal& jm = arr("meshes");
        for (size_t mi = 0; mi < jm.size(); ++mi) {
            const JVal* prims = jm.at(mi).get("primitives");
            std::vector<float> P, N, UV, T;
            std::vector<uint32_t> I, M;
            bool any_tan = false;
            uint32_t section = 0;
            for (size_t pi = 0; prims && pi < prims->size(); ++pi) {
                const JVal& pr = prims->at(pi);
                if (pr.int_or("mode", 4) != 4) continue;
                const JVal* at = pr.get("attributes");
                const JVal* pa = at ? at->get("POSITION") : nullptr;
                if (!pa) continue;
                const std::vector<float> pp = floats(int(pa->num), 3);
                const size_t nv = pp.size() / 3;
                std::vector<uint32_t> ii;
                if (const JVal* ia = pr.get("indices")) {
                    ii = indices(int(ia->num));
                } else {
                    ii.resize(nv);
                    for (size_t k = 0; k < nv; ++k) ii[k] = uint32_t(k);
                }
                ii.resize(ii.size() / 3 * 3);
                if (ii.empty()) continue;
                for (const uint32_t x : ii)
                    if (x >= nv) fail(GK_ERR_ARG, "glTF: index out of range");
                std::vector<float> nn = at->get("NORMAL") ? floats(int(at->get("NORMAL")->num), 3) : std::vector<float>();
                if (nn.size() != pp.size()) {  // no normals: area-weighted face normals
                    nn.assign(pp.size(), 0.f);
                    for (size_t t = 0; t < ii.size(); t += 3) {
                        const float* a = &pp[3 * ii[t]]; const float* b = &pp[3 * ii[t + 1]]; const float* c = &pp[3 * ii[t + 2]];
                        const float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
                        const float f[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
                        for (int k = 0; k < 3; ++k)
                            for (int j = 0; j < 3; ++j) nn[3 * ii[t + k] + j] += f[j];
                    }
                    for (size_t v = 0; v < nv; ++v) {
                        float* q = &nn[3 * v];
                        const float l = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
                        if (l > 0.f) { q[0] /= l; q[1] /= l; q[2] /= l; } else { q[1] = 1.f; }
                    }
                }
                std::vector<float> tt = at->get("TEXCOORD_0") ? floats(int(at->get("TEXCOORD_0")->num), 2) : std::vector<float>();
                if (tt.size() != 2 * nv) tt.assign(2 * nv, 0.f);
                std::vector<float> tg;
                if (at->get("TANGENT")) {
                    tg = floats(int(at->get("TANGENT")->num), 4);
                    any_tan = true;
                }
                if (tg.size() != 4 * nv) tg.assign(4 * nv, 0.f);
                const uint32_t base = uint32_t(P.size() / 3);
                P.insert(P.end(), pp.begin(), pp.end());
                N.insert(N.end(), nn.begin(), nn.end());
                UV.insert(UV.end(), tt.begin(), tt.end());
                T.insert(T.end(), tg.begin(), tg.end());
                for (const uint32_t x : ii) I.push_back(x + base);
                M.insert(M.end(), ii.size() / 3, section);
                section++;
            }
            if (!any_tan) gen_tangents(P, N, UV, I, T);
            pos.push_back(std::move(P));
            nrm.push_back(std::move(N));
            uv.push_back(std::move(UV));
            tan.push_back(std::move(T));
            idx.push_back(std::move(I));
            tmat.push_back(std::move(M));
        }
        for (size_t mi = 0; mi < pos.size(); ++mi) {
            gk_mesh g{};
            g.positions = pos[mi].data();
            g.normals = nrm[mi].data();
            g.texcoords = uv[mi].data();
            g.tangents = tan[mi].data();
            g.indices = idx[mi].data();
            g.tri_material = tmat[mi].data();
            g.num_vertices = uint32_t(pos[mi].size() / 3);
            g.num_triangles = uint32_t(idx[mi].size() / 3);
            meshes.push_back(g);
        }
        // scene 0 (FSceneLoader.cpp:510-566)
        const JVal& scenes = arr("scenes");
        if (scenes.size() > 0) {
            const JVal& root = scenes.at(0);
            if (const JVal* ex = root.get("extras")) {
                if (ex->get("SkyIdx")) { view.has_sky = 1; view.sky_idx = ex->int_or("SkyIdx", 0); }
                if (ex->get("SkyIntensity")) { view.has_sky = 1; view.sky_intensity = float(ex->num_or("SkyIntensity", 1)); }
                if (ex->get("SkyRotation")) { view.has_sky = 1; view.sky_rotation = float(ex->num_or("SkyRotation", 0)); }
                if (ex->get("SunIntensity")) { view.has_sun = 1; view.sun_intensity = float(ex->num_or("SunIntensity", 1)); }
                if (ex->get("WithSun")) view.has_sun = ex->int_or("WithSun", 0) != 0 ? 1u : 0u;
                if (ex->get("SunRotation")) view.sun_rotation = float(ex->num_or("SunRotation", 0));
                if (ex->get("NoSky")) view.has_sky = 0;
            }
            Mat4 id{};
            for (int k = 0; k < 4; ++k) id.m[k * 5] = 1.f;
            uint32_t node_count = 0;
            if (const JVal* ns = root.get("nodes"))
                for (size_t i = 0; i < ns->size(); ++i) node(int(ns->at(i).num), id, node_count, 0);
        }
    }

    // ParseGltfNode (FSceneLoader.cpp:102-188), depth first
    void node(int ni, const Mat4& parent, uint32_t& count, int depth) {
        if (depth > 64) fail(GK_ERR_ARG, "glTF: node hierarchy too deep");
        const JVal& n = arr("nodes").at(size_t(ni));
        float t[3] = {0, 0, 0}, sc[3] = {1, 1, 1}, q[4] = {0, 0, 0, 1};  // q = x, y, z, w
        if (const JVal* v = n.get("translation")) for (size_t k = 0; k < 3 && k < v->size(); ++k) t[k] = float(v->at(k).num);
        if (const JVal* v = n.get("scale")) for (size_t k = 0; k < 3 && k < v->size(); ++k) sc[k] = float(v->at(k).num);
        if (const JVal* v = n.get("rotation")) for (size_t k = 0; k < 4 && k < v->size(); ++k) q[k] = float(v->at(k).num);
        const float ql = std::sqrt(q[3] * q[3] + q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        if (ql > 0.f) for (float& c : q) c /= ql;
        const float x = q[0], y = q[1], z = q[2], w = q[3];
        Mat4 local{};
        // T * R * S, R the rotation matrix of the unit quaternion (x, y, z, w)
        const float R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y),
                            2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x),
                            2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)};
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) local.m[c * 4 + r] = R[c * 3 + r] * sc[c];
        local.m[12] = t[0]; local.m[13] = t[1]; local.m[14] = t[2]; local.m[15] = 1.f;
        const Mat4 world = mul(parent, local);
        const uint32_t self = count++;
        const int mesh = n.int_or("mesh", -1);
        uint32_t primary_mat = 0;
        if (mesh >= 0) {
            if (size_t(mesh) >= meshes.size()) fail(GK_ERR_ARG, "glTF: node mesh index");
            gk_instance in{};
            std::memcpy(in.world, world.m, 64);
            std::memcpy(in.prev_world, world.m, 64);
            in.instance_id = self;
            in.mesh_id = uint32_t(mesh);
            in.visible = 1;
            in.use_remap = 1;
            const JVal* prims = arr("meshes").at(size_t(mesh)).get("primitives");
            const size_t np = prims ? prims->size() : 0;
            if (np > 16) fail(GK_ERR_LIMIT, "glTF: more than 16 primitives in a mesh (the node material map has 16 slots)");
            for (size_t i = 0; i < np; ++i) {
                const int pm = prims->at(i).int_or("material", -1);
                in.mat_id[i] = uint32_t(std::max(0, pm));
                primary_mat = uint32_t(pm);
            }
            instances.push_back(in);
        } else if (n.get("camera")) {
            // the node's own transform (the parent is not applied), fov 40 (FSceneLoader.cpp:127-139)
            const float o4[4] = {0, 0, 0, 1}, f4[4] = {0, 0, -1, 0};
            float eye[4], fwd[4];
            xform(local, o4, eye);
            xform(local, f4, fwd);
            if (view.num_cameras == 0) {
                gk_camera_desc cam{};
                for (int k = 0; k < 3; ++k) { cam.eye[k] = eye[k]; cam.target[k] = eye[k] + fwd[k]; }
                cam.up[1] = 1.f;
                cam.fov_y_degrees = 40.f;
                cam.focus_distance = 10.f;
                view.camera = cam;
            }
            view.num_cameras++;
        }
        const JVal* ex = n.get("extras");
        if (ex && ex->get("arealight")) {
            // the quad from -1 to 1 in the node's transform at creation, before its parent is
            // set (FSceneLoader.cpp:152-172)
            const float l0[4] = {-1, 0, -1, 1}, l1[4] = {-1, 0, 1, 1}, l3[4] = {1, 0, -1, 1}, up[4] = {0, 1, 0, 0};
            gk_light L{};
            xform(local, l0, L.p0);
            xform(local, l1, L.p1);
            xform(local, l3, L.p3);
            float d[4];
            xform(local, up, d);
            const float dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            for (int k = 0; k < 3; ++k) L.normal_area[k] = dl > 0.f ? d[k] / dl : 0.f;
            const float a[3] = {L.p1[0] - L.p0[0], L.p1[1] - L.p0[1], L.p1[2] - L.p0[2]};
            const float b[3] = {L.p3[0] - L.p0[0], L.p3[1] - L.p0[1], L.p3[2] - L.p0[2]};
            const float c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
            L.normal_area[3] = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) / 2.0f;
            L.light_mat_idx = primary_mat;
            lights.push_back(L);
        }
        if (const JVal* ch = n.get("children"))
            for (size_t i = 0; i < ch->size(); ++i) node(int(ch->at(i).num), world, count, depth + 1);
    }

    // Per-vertex tangents from texture-coordinate derivatives (Lengyel's method: the
    // triangle tangents summed per vertex, made orthogonal to the normal, handedness from
    // the bitangent), used when a mesh has no TANGENT attribute; the reference generates
    // MikkTSpace tangents there (FSceneLoader.cpp:200-207).
    static void gen_tangents(const std::vector<float>& P, const std::vector<float>& N, const std::vector<float>& UV,
                             const std::vector<uint32_t>& I, std::vector<float>& T) {
        const size_t nv = P.size() / 3;
        std::vector<double> tu(3 * nv, 0.0), tb(3 * nv, 0.0);
        for (size_t t = 0; t + 2 < I.size(); t += 3) {
            const uint32_t a = I[t], b = I[t + 1], c = I[t + 2];
            double e1[3], e2[3];
            for (int k = 0; k < 3; ++k) { e1[k] = P[3 * b + k] - P[3 * a + k]; e2[k] = P[3 * c + k] - P[3 * a + k]; }
            const double du1 = UV[2 * b] - UV[2 * a], dv1 = UV[2 * b + 1] - UV[2 * a + 1];
            const double du2 = UV[2 * c] - UV[2 * a], dv2 = UV[2 * c + 1] - UV[2 * a + 1];
            const double det = du1 * dv2 - du2 * dv1;
            if (std::fabs(det) < 1e-20) continue;
            const double r = 1.0 / det;
            for (const uint32_t v : {a, b, c})
                for (int k = 0; k < 3; ++k) {
                    tu[3 * v + k] += (e1[k] * dv2 - e2[k] * dv1) * r;
                    tb[3 * v + k] += (e2[k] * du1 - e1[k] * du2) * r;
                }
        }
        T.assign(4 * nv, 0.f);
        for (size_t v = 0; v < nv; ++v) {
            const double n[3] = {N[3 * v], N[3 * v + 1], N[3 * v + 2]};
            double t[3] = {tu[3 * v], tu[3 * v + 1], tu[3 * v + 2]};
            const double d = n[0] * t[0] + n[1] * t[1] + n[2] * t[2];
            for (int k = 0; k < 3; ++k) t[k] -= n[k] * d;
            double l = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
            if (l < 1e-12) {  // no usable derivative: any direction orthogonal to n
                const double ax[3] = {std::fabs(n[0]) < 0.9 ? 1.0 : 0.0, std::fabs(n[0]) < 0.9 ? 0.0 : 1.0, 0.0};
                const double dd = n[0] * ax[0] + n[1] * ax[1] + n[2] * ax[2];
                for (int k = 0; k < 3; ++k) t[k] = ax[k] - n[k] * dd;
                l = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
            }
            for (int k = 0; k < 3; ++k) T[4 * v + k] = float(t[k] / l);
            const double c[3] = {n[1] * t[2] - n[2] * t[1], n[2] * t[0] - n[0] * t[2], n[0] * t[1] - n[1] * t[0]};
            const double h = c[0] * tb[3 * v] + c[1] * tb[3 * v + 1] + c[2] * tb[3 * v + 2];
            T[4 * v + 3] = h < 0.0 ? -1.f : 1.f;
        }
    }

    void fill_view() {
        view.meshes = meshes.data();
        view.num_meshes = uint32_t(meshes.size());
        view.materials = materials.data();
        view.num_materials = uint32_t(materials.size());
        view.instances = instances.data();
        view.num_instances = uint32_t(instances.size());
        view.lights = lights.data();
        view.num_lights = uint32_t(lights.size());
        view.textures = textures.data();
        view.num_textures = uint32_t(textures.size());
    }
};

constexpr uint32_t kGlbMagic = 0x46546C67u, kChunkJson = 0x4E4F534Au, kChunkBin = 0x004E4942u;

uint32_t le32(const uint8_t* b) { return uint32_t(b[0]) | (uint32_t(b[1]) << 8) | (uint32_t(b[2]) << 16) | (uint32_t(b[3]) << 24); }

std::unique_ptr<Scene> parse(const std::vector<uint8_t>& bytes, const std::string& dir) {
    auto sc = std::make_unique<Scene>();
    sc->dir = dir;
    std::vector<uint8_t> bin;
    bool has_bin = false;
    const char* js = nullptr;
    size_t js_len = 0;
    if (bytes.size() >= 12 && le32(bytes.data()) == kGlbMagic) {
        if (le32(&bytes[4]) != 2) fail(GK_ERR_ARG, "glTF: GLB version");
        const size_t total = std::min<size_t>(le32(&bytes[8]), bytes.size());
        size_t p = 12;
        while (p + 8 <= total) {
            const uint32_t len = le32(&bytes[p]), type = le32(&bytes[p + 4]);
            if (p + 8 + size_t(len) > total) fail(GK_ERR_ARG, "glTF: truncated GLB chunk");
            if (type == kChunkJson && !js) { js = reinterpret_cast<const char*>(&bytes[p + 8]); js_len = len; }
            else if (type == kChunkBin && !has_bin) { bin.assign(bytes.begin() + p + 8, bytes.begin() + p + 8 + len); has_bin = true; }
            p += 8 + size_t(len);
        }
        if (!js) fail(GK_ERR_ARG, "glTF: GLB without JSON chunk");
    } else {
        js = reinterpret_cast<const char*>(bytes.data());
        js_len = bytes.size();
    }
    JParser jp{js, js + js_len};
    sc->doc = jp.value();
    if (sc->doc.t != JVal::Obj) fail(GK_ERR_ARG, "glTF: top level is not an object");
    const JVal& bufs = sc->arr("buffers");
    for (size_t i = 0; i < bufs.size(); ++i) {
        const JVal* uri = bufs.at(i).get("uri");
        if (uri) sc->buffers.push_back(uri_bytes(uri->str, dir));
        else if (i == 0 && has_bin) sc->buffers.push_back(bin);
        else fail(GK_ERR_ARG, "glTF: buffer without data");
    }
    sc->build();
    sc->fill_view();
    return sc;
}

}  // namespace

struct gk_gltf {
    std::unique_ptr<Scene> s;
};

extern "C" {

const char* gk_gltf_error(void) { return g_err.c_str(); }

int gk_gltf_load(const char* path, gk_gltf** out) {
    if (!path || !out) return GK_ERR_ARG;
    *out = nullptr;
    g_err.clear();
    try {
        const std::string p(path);
        const size_t slash = p.find_last_of('/');
        const std::string dir = slash == std::string::npos ? std::string() : p.substr(0, slash);
        auto sc = parse(read_file(p), dir);
        *out = new gk_gltf{std::move(sc)};
        return GK_OK;
    } catch (const Fail& f) {
        return f.code;
    } catch (const std::bad_alloc&) {
        g_err = "glTF: out of host memory";
        return GK_ERR_OOM;
    }
}

int gk_gltf_load_memory(const void* glb, size_t bytes, gk_gltf** out) {
    if (!glb || !out) return GK_ERR_ARG;
    *out = nullptr;
    g_err.clear();
    try {
        const uint8_t* b = static_cast<const uint8_t*>(glb);
        auto sc = parse(std::vector<uint8_t>(b, b + bytes), std::string());
        *out = new gk_gltf{std::move(sc)};
        return GK_OK;
    } catch (const Fail& f) {
        return f.code;
    } catch (const std::bad_alloc&) {
        g_err = "glTF: out of host memory";
        return GK_ERR_OOM;
    }
}

int gk_gltf_get_view(gk_gltf* scene, gk_gltf_view* out) {
    if (!scene || !out) return GK_ERR_ARG;
    *out = scene->s->view;
    return GK_OK;
}

int gk_gltf_upload(gk_ctx* ctx, gk_gltf* scene) {
    if (!ctx || !scene) return GK_ERR_ARG;
    const gk_gltf_view& v = scene->s->view;
    int st = gk_scene_upload(ctx, v.meshes, v.num_meshes, v.materials, v.num_materials,
                             v.num_lights ? v.lights : nullptr, v.num_lights, v.instances, v.num_instances);
    if (st != GK_OK) return st;
    return gk_textures_upload(ctx, v.num_textures ? v.textures : nullptr, v.num_textures);
}

int gk_gltf_free(gk_gltf* scene) {
    delete scene;
    return GK_OK;
}

}  // extern "C"
