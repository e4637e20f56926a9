// Scene loading (reference: core/src/Scene.cpp).  OBJ parsing follows what
// TriangleMesh::LoadMesh keeps from tinyobj's output (Scene.cpp:83-179):
// polygons are fan-triangulated, every triangle gets three fresh vertices,
// a missing normal becomes the flat geometric normal and a missing uv
// becomes (0,0),(1,0),(0,1).
#include "Scene.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>
#include <stdexcept>

// Inside the reference's tree (common/ on the include path, CMakeLists.txt:37,
// 64-68) the facade decodes images with the reference's own vendored
// stb_image, exactly as its Texture and HdrTexture do (Scene.cpp:181-206):
// PNG, JPG, BMP, PNM and Radiance HDR, rows flipped on load.  Built on its
// own, it reads binary PPM, uncompressed BMP, Radiance HDR and PFM with the
// readers below.
#if __has_include("3rdParty/stb_image.h")
#define STB_IMAGE_IMPLEMENTATION
#include "3rdParty/stb_image.h"
#define OXR_FACADE_STB 1
#endif

namespace {

struct ObjIndex { int v = -1, t = -1, n = -1; };

int fix_index(int i, size_t count) {
    if (i > 0) return i - 1;
    if (i < 0) return (int)count + i;
    return -1;
}

ObjIndex parse_vertex(const std::string& tok, size_t nv, size_t nt, size_t nn) {
    ObjIndex r;
    int parts[3] = {0, 0, 0};
    bool have[3] = {false, false, false};
    size_t pos = 0;
    for (int k = 0; k < 3 && pos <= tok.size(); k++) {
        size_t slash = tok.find('/', pos);
        std::string s = tok.substr(pos, slash == std::string::npos ? std::string::npos : slash - pos);
        if (!s.empty()) { parts[k] = std::stoi(s); have[k] = true; }
        if (slash == std::string::npos) break;
        pos = slash + 1;
    }
    if (have[0]) r.v = fix_index(parts[0], nv);
    if (have[1]) r.t = fix_index(parts[1], nt);
    if (have[2]) r.n = fix_index(parts[2], nn);
    return r;
}

}  // namespace

void TriangleMesh::LoadMesh(const std::string& objFile, const Material& mat, int out_med, int in_med) {
    std::ifstream in(objFile);
    if (!in) throw std::runtime_error("Could not read OBJ model from " + objFile);
    material = mat;
    in_medium = in_med;
    out_medium = out_med;
    std::vector<vec3f> V, N;
    std::vector<vec2f> T;
    std::string line;
    size_t faces = 0;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string tag;
        if (!(ss >> tag)) continue;
        if (tag == "v") {
            vec3f p;
            ss >> p.x >> p.y >> p.z;
            V.push_back(p);
        } else if (tag == "vn") {
            vec3f p;
            ss >> p.x >> p.y >> p.z;
            N.push_back(p);
        } else if (tag == "vt") {
            vec2f p;
            ss >> p.x >> p.y;
            T.push_back(p);
        } else if (tag == "f") {
            std::vector<ObjIndex> poly;
            std::string tok;
            while (ss >> tok) poly.push_back(parse_vertex(tok, V.size(), T.size(), N.size()));
            for (size_t k = 2; k < poly.size(); k++) {
                const ObjIndex* tri[3] = {&poly[0], &poly[k - 1], &poly[k]};
                vec3f p[3], n[3];
                vec2f t[3];
                bool has_n = true, has_t = true;
                for (int i = 0; i < 3; i++) {
                    if (tri[i]->v < 0 || tri[i]->v >= (int)V.size())
                        throw std::runtime_error("OBJ " + objFile + ": vertex index out of range");
                    p[i] = V[tri[i]->v];
                    if (tri[i]->n >= 0 && tri[i]->n < (int)N.size()) n[i] = normalize(N[tri[i]->n]);
                    else has_n = false;
                    if (tri[i]->t >= 0 && tri[i]->t < (int)T.size()) t[i] = T[tri[i]->t];
                    else has_t = false;
                }
                if (!has_n) {  // Scene.cpp:150-158
                    vec3f e1 = normalize(p[1] - p[0]), e2 = normalize(p[2] - p[0]);
                    n[0] = n[1] = n[2] = normalize(cross(e1, e2));
                }
                if (!has_t) {  // Scene.cpp:160-165
                    t[0] = vec2f(0.0f);
                    t[1] = vec2f(1.0f, 0.0f);
                    t[2] = vec2f(0.0f, 1.0f);
                }
                int base = (int)vertex.size();
                for (int i = 0; i < 3; i++) {
                    vertex.push_back(p[i]);
                    normal.push_back(n[i]);
                    texcoord.push_back(t[i]);
                }
                index.push_back(vec3i(base, base + 1, base + 2));
                faces++;
            }
        }
    }
    if (faces == 0) throw std::runtime_error("Could not read OBJ model from " + objFile + " : no faces");
}

namespace {

uint32_t le32(const uint8_t* p) { return p[0] | (p[1] << 8) | (p[2] << 16) | ((uint32_t)p[3] << 24); }

// Uncompressed 24- or 32-bit BMP.  Rows end up bottom first, as stbi's
// flipped load leaves them.
void read_bmp(const std::string& fileName, std::ifstream& in, std::vector<uint8_t>& pixel, vec2i& res) {
    std::vector<uint8_t> f((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    auto bad = [&](const char* why) { throw std::runtime_error("texture " + fileName + ": " + why); };
    if (f.size() < 54 || f[0] != 'B' || f[1] != 'M') bad("not a BMP file");
    const uint32_t off = le32(&f[10]);
    const int32_t w = (int32_t)le32(&f[18]), h_signed = (int32_t)le32(&f[22]);
    const int bpp = f[28] | (f[29] << 8);
    const uint32_t compression = le32(&f[30]);
    if (w <= 0 || h_signed == 0 || w > (1 << 15) || h_signed > (1 << 15) || h_signed < -(1 << 15))
        bad("bad BMP size");
    if ((bpp != 24 && bpp != 32) || (compression != 0 && !(compression == 3 && bpp == 32)))
        bad("only uncompressed 24- and 32-bit BMP files are supported");
    const int h = h_signed < 0 ? -h_signed : h_signed;
    const size_t stride = ((size_t)w * (bpp / 8) + 3) & ~(size_t)3;
    if (off > f.size() || stride * h > f.size() - off) bad("truncated BMP");
    pixel.resize((size_t)w * h * 4);
    for (int y = 0; y < h; y++) {
        const int row = h_signed > 0 ? y : h - 1 - y;  // bottom-up files already start at the bottom
        const uint8_t* s = &f[off + stride * y];
        for (int x = 0; x < w; x++) {
            uint8_t* d = &pixel[4 * ((size_t)row * w + x)];
            const uint8_t* px = s + (size_t)x * (bpp / 8);
            d[0] = px[2]; d[1] = px[1]; d[2] = px[0]; d[3] = 255;  // BGR(A) -> RGBA
        }
    }
    res = vec2i(w, h);
}

}  // namespace

Texture::Texture(const std::string& fileName) {
#ifdef OXR_FACADE_STB
    {
        int w = 0, h = 0, comp_in = 0;
        stbi_set_flip_vertically_on_load(true);
        uint8_t* image = stbi_load(fileName.c_str(), &w, &h, &comp_in, STBI_rgb_alpha);
        if (!image)
            throw std::runtime_error("Could not read texture " + fileName + ": " + stbi_failure_reason());
        pixel.assign(image, image + (size_t)w * h * 4);
        stbi_image_free(image);
        resolution = vec2i(w, h);
        return;
    }
#endif
    std::ifstream in(fileName, std::ios::binary);
    if (!in) throw std::runtime_error("Could not read texture " + fileName);
    if (in.peek() == 'B') {
        read_bmp(fileName, in, pixel, resolution);
        return;
    }
    std::string magic;
    in >> magic;
    if (magic != "P6") throw std::runtime_error("texture " + fileName + ": only binary PPM (P6) and BMP are supported");
    int w, h, maxv;
    in >> w >> h >> maxv;
    in.get();
    if (w <= 0 || h <= 0 || maxv != 255) throw std::runtime_error("texture " + fileName + ": bad header");
    std::vector<uint8_t> rgb((size_t)w * h * 3);
    in.read((char*)rgb.data(), rgb.size());
    if (!in) throw std::runtime_error("texture " + fileName + ": truncated");
    pixel.resize((size_t)w * h * 4);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const uint8_t* s = &rgb[3 * ((size_t)(h - 1 - y) * w + x)];  // flipped like stbi
            uint8_t* d = &pixel[4 * ((size_t)y * w + x)];
            d[0] = s[0]; d[1] = s[1]; d[2] = s[2]; d[3] = 255;
        }
    resolution = vec2i(w, h);
}

namespace {

void rgbe_to_float(const uint8_t* e, float* out) {
    if (e[3] == 0) { out[0] = out[1] = out[2] = 0.0f; return; }
    float f = std::ldexp(1.0f, (int)e[3] - (128 + 8));
    out[0] = e[0] * f; out[1] = e[1] * f; out[2] = e[2] * f;
}

bool read_rgbe_scanline(std::istream& in, int w, std::vector<uint8_t>& line) {
    line.assign((size_t)4 * w, 0);
    uint8_t head[4];
    if (!in.read((char*)head, 4)) return false;
    if (w < 8 || w > 0x7fff || head[0] != 2 || head[1] != 2 || (head[2] & 0x80)) {
        // flat scanline
        std::memcpy(&line[0], head, 4);
        return (bool)in.read((char*)&line[4], (size_t)4 * (w - 1));
    }
    if (((head[2] << 8) | head[3]) != w) return false;
    for (int c = 0; c < 4; c++) {
        int x = 0;
        while (x < w) {
            int count = in.get();
            if (count < 0) return false;
            if (count > 128) {
                count -= 128;
                int v = in.get();
                if (v < 0 || x + count > w) return false;
                for (int k = 0; k < count; k++) line[4 * (x++) + c] = (uint8_t)v;
            } else {
                if (count == 0 || x + count > w) return false;
                for (int k = 0; k < count; k++) {
                    int v = in.get();
                    if (v < 0) return false;
                    line[4 * (x++) + c] = (uint8_t)v;
                }
            }
        }
    }
    return true;
}

}  // namespace

HdrTexture::HdrTexture(int w, int h, std::vector<float> rgba) : hdr(std::move(rgba)), width(w), height(h) {
    if (w <= 0 || h <= 0 || hdr.size() != (size_t)w * h * 4) throw std::runtime_error("HdrTexture: bad size");
}

HdrTexture::HdrTexture(const std::string& fileName) {
    std::ifstream in(fileName, std::ios::binary);
    if (!in) {
        // The reference attaches an environment without data when the file
        // cannot be read (Scene.cpp:197-206); here that is no environment
        // (misses are black) and a warning, so main.cpp, whose HDR is not in
        // the reference's models/ folder, still runs.
        std::fprintf(stderr, "warning: could not read HDR %s: no environment\n", fileName.c_str());
        return;
    }
#ifdef OXR_FACADE_STB
    {
        in.close();
        int w = 0, h = 0, comp_in = 0;
        stbi_set_flip_vertically_on_load(true);
        float* data = stbi_loadf(fileName.c_str(), &w, &h, &comp_in, STBI_rgb_alpha);
        if (!data) {
            std::fprintf(stderr, "warning: could not decode HDR %s (%s): no environment\n", fileName.c_str(),
                         stbi_failure_reason());
            return;
        }
        hdr.assign(data, data + (size_t)w * h * 4);
        stbi_image_free(data);
        width = w;
        height = h;
        return;
    }
#endif
    std::string first;
    std::getline(in, first);
    if (first == "PF") {  // PFM, bottom-to-top rows already
        int w, h;
        float scale;
        in >> w >> h >> scale;
        in.get();
        if (w <= 0 || h <= 0) throw std::runtime_error("HDR " + fileName + ": bad PFM header");
        std::vector<float> rgb((size_t)w * h * 3);
        in.read((char*)rgb.data(), rgb.size() * 4);
        if (!in) throw std::runtime_error("HDR " + fileName + ": truncated");
        width = w; height = h;
        hdr.resize((size_t)w * h * 4);
        for (size_t i = 0; i < (size_t)w * h; i++) {
            hdr[4 * i] = rgb[3 * i]; hdr[4 * i + 1] = rgb[3 * i + 1]; hdr[4 * i + 2] = rgb[3 * i + 2]; hdr[4 * i + 3] = 1.0f;
        }
        return;
    }
    if (first.rfind("#?", 0) != 0) throw std::runtime_error("HDR " + fileName + ": not a Radiance file");
    std::string line;
    while (std::getline(in, line) && !line.empty()) {
        if (line.rfind("FORMAT=", 0) == 0 && line != "FORMAT=32-bit_rle_rgbe")
            throw std::runtime_error("HDR " + fileName + ": unsupported " + line);
    }
    std::getline(in, line);
    char ya[3], xa[3];
    int h = 0, w = 0;
    if (std::sscanf(line.c_str(), "%2s %d %2s %d", ya, &h, xa, &w) != 4 || std::string(ya) != "-Y" ||
        std::string(xa) != "+X" || w <= 0 || h <= 0)
        throw std::runtime_error("HDR " + fileName + ": unsupported resolution line '" + line + "'");
    width = w;
    height = h;
    hdr.resize((size_t)w * h * 4);
    std::vector<uint8_t> sl;
    for (int y = 0; y < h; y++) {
        if (!read_rgbe_scanline(in, w, sl)) throw std::runtime_error("HDR " + fileName + ": bad scanline");
        int row = h - 1 - y;  // file is top to bottom; flip like stbi_set_flip_vertically_on_load(true)
        for (int x = 0; x < w; x++) {
            float* d = &hdr[4 * ((size_t)row * w + x)];
            rgbe_to_float(&sl[4 * x], d);
            d[3] = 1.0f;
        }
    }
}

Scene::~Scene() {
    for (auto m : meshes) delete m;
    for (auto t : textures) delete t;
    delete env;
}

void Scene::AddMesh(const std::string& objFile, Material& material, const TextureFile& tf, int out_med, int in_med) {
    auto add = [&](const std::string& f, int& id) {
        if (!f.empty()) {
            id = (int)textures.size();
            AddTexture(f);
        }
    };
    add(tf.albedoFile, material.albedoTextureID);
    add(tf.anisotropyFile, material.anisotropyTextureID);
    add(tf.roughnessFile, material.roughnessTextureID);
    add(tf.specularFile, material.specularTextureID);
    add(tf.metallicFile, material.metallicTextureID);
    add(tf.coat_roughness_uFile, material.coat_roughness_uTextureID);
    add(tf.coat_roughness_vFile, material.coat_roughness_vTextureID);
    add(tf.normalFile, material.normalTextureID);
    TriangleMesh* mesh = new TriangleMesh();
    try {
        mesh->LoadMesh(objFile, material, out_med, in_med);
    } catch (...) {
        delete mesh;
        throw;
    }
    meshes.push_back(mesh);
    for (auto& v : mesh->vertex) bounds.extend(v);
}

void Scene::AddTriangles(const std::vector<float>& positions, const std::vector<float>& normals,
                         const Material& material, int out_med, int in_med) {
    if (positions.size() % 9 != 0 || normals.size() != positions.size())
        throw std::runtime_error("AddTriangles: 9 floats of positions and normals per triangle");
    TriangleMesh* mesh = new TriangleMesh();
    mesh->material = material;
    mesh->in_medium = in_med;
    mesh->out_medium = out_med;
    for (size_t i = 0; i < positions.size(); i += 3) {
        mesh->vertex.push_back(vec3f(positions[i], positions[i + 1], positions[i + 2]));
        mesh->normal.push_back(vec3f(normals[i], normals[i + 1], normals[i + 2]));
        const int corner = (int)(i / 3) % 3;  // (0,0), (1,0), (0,1) as LoadMesh leaves untextured faces
        mesh->texcoord.push_back(vec2f(corner == 1 ? 1.0f : 0.0f, corner == 2 ? 1.0f : 0.0f));
    }
    for (int t = 0; t < (int)positions.size() / 9; t++) mesh->index.push_back(vec3i(3 * t, 3 * t + 1, 3 * t + 2));
    meshes.push_back(mesh);
    for (auto& v : mesh->vertex) bounds.extend(v);
}

void Scene::AddMesh(TriangleMesh* mesh) {
    meshes.push_back(mesh);
    for (auto& v : mesh->vertex) bounds.extend(v);
}

void Scene::AddTexture(const std::string& fileName) { textures.push_back(new Texture(fileName)); }
void Scene::AddLight(const Light& l) { lights.push_back(l); }
void Scene::AddMedium(const Medium& m) { mediums.push_back(m); }
void Scene::AddEnv(const std::string& fileName) {
    HdrTexture* e = new HdrTexture(fileName);
    delete env;
    env = e;
}
void Scene::AddEnv(HdrTexture* e) {
    delete env;
    env = e;
}
