// rb_io.cpp -- scene ingestion: Wavefront OBJ meshes and the reference's XML
// scene description.
//
// ObjFileReader::read_file (src/io/objfilereader.cpp:117-217): `o` starts an
// object, `v` / `vn` / `vt` lines add data, `f` lines add faces with 1-based
// indices counted over the whole file; the reference's Scene::add_mesh
// (src/core/scene.cpp:83-113) concatenates the objects' vertices in file
// order without re-basing the indices.  This reader keeps that result:
// vertices in file order, triangles in file order, file-global indices.
// Differences (each one a case where the reference loses or corrupts data):
//  * faces of more than three corners are fanned into triangles (the
//    reference keeps the first three corners);
//  * negative (relative) indices are resolved (the reference subtracts 1
//    from them);
//  * `#` starts a comment anywhere on a line (the reference drops any line
//    containing `#`, including `v 1 2 3 # note`);
//  * data before the first `o` goes to an unnamed object (the reference
//    dereferences a null mesh).
//
// CustomFileReader::process / get_camera (src/io/customfilereader.cpp:47-340):
// <scene> with <files><count>1</count><file>mesh.obj</file></files>,
// <material> entries (<name>, <bsdf><layers>n</layers><bxdf><type>..</type>
// <reflectance>{r,g,b}</reflectance>...</bxdf></bsdf>), <object> entries
// assigning a material by name to an OBJ object, <lightobjects><light> entries
// (<name> of an OBJ object, <Le>{r,g,b}</Le>) making that object a diffuse
// emitter, and <camera> (<position>, <lookat>, <up>, <fov>).  Materials are
// numbered in file order, then one emitter material per light, as the
// reference numbers them.  Lambertian, mirror, glass, orennayar and (Blinn)
// microfacet bxdfs map to the renderer's materials; a material with 2..4
// <bxdf> layers becomes an RB_MAT_LAYERED material whose layers are added
// after the emitters.  The camera's <fov> is used when present
// (the reference hard-codes 90 degrees, customfilereader.cpp:331-337).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>

#include "rb_internal.h"

namespace rb {
namespace {

// ------------------------------------------------------------------ XML --

struct XNode {
    std::string name, text;
    std::vector<XNode> kids;
    const XNode* child(const char* n) const {
        for (const XNode& k : kids)
            if (k.name == n) return &k;
        return nullptr;
    }
};

struct XmlParser {
    const std::string& s;
    size_t i = 0;
    std::string err;
    explicit XmlParser(const std::string& src) : s(src) {}

    bool starts(const char* t) const { return s.compare(i, std::strlen(t), t) == 0; }
    void skip_ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\r' || s[i] == '\n')) i++;
    }
    bool skip_past(const char* t) {
        size_t p = s.find(t, i);
        if (p == std::string::npos) return false;
        i = p + std::strlen(t);
        return true;
    }
    // comments, declarations, processing instructions
    bool skip_misc() {
        for (;;) {
            skip_ws();
            if (starts("<!--")) {
                if (!skip_past("-->")) return fail_at("unterminated comment");
            } else if (starts("<?")) {
                if (!skip_past("?>")) return fail_at("unterminated declaration");
            } else if (starts("<!")) {
                if (!skip_past(">")) return fail_at("unterminated markup");
            } else {
                return true;
            }
        }
    }
    bool fail_at(const char* what) {
        size_t line = 1;
        for (size_t k = 0; k < i && k < s.size(); k++) line += s[k] == '\n';
        char buf[160];
        std::snprintf(buf, sizeof buf, "line %zu: %s", line, what);
        err = buf;
        return false;
    }
    static std::string decode(const std::string& t) {
        std::string o;
        for (size_t k = 0; k < t.size(); k++) {
            if (t[k] == '&') {
                static const char* ent[][2] = {{"&lt;", "<"}, {"&gt;", ">"}, {"&amp;", "&"}, {"&quot;", "\""}, {"&apos;", "'"}};
                bool done = false;
                for (auto& e : ent)
                    if (t.compare(k, std::strlen(e[0]), e[0]) == 0) {
                        o += e[1];
                        k += std::strlen(e[0]) - 1;
                        done = true;
                        break;
                    }
                if (!done) o += '&';
            } else {
                o += t[k];
            }
        }
        return o;
    }
    static std::string trim(const std::string& t) {
        size_t a = t.find_first_not_of(" \t\r\n"), b = t.find_last_not_of(" \t\r\n");
        return a == std::string::npos ? std::string() : t.substr(a, b - a + 1);
    }
    bool element(XNode& n, int depth) {
        if (depth > 64) return fail_at("elements nested too deeply");
        if (i >= s.size() || s[i] != '<') return fail_at("expected an element");
        i++;
        size_t a = i;
        while (i < s.size() && !std::strchr(" \t\r\n/>", s[i])) i++;
        n.name = s.substr(a, i - a);
        if (n.name.empty()) return fail_at("empty element name");
        size_t close = s.find('>', i);  // attributes are not used by the format
        if (close == std::string::npos) return fail_at("unterminated tag");
        bool self = close > 0 && s[close - 1] == '/';
        i = close + 1;
        if (self) return true;
        std::string text;
        for (;;) {
            size_t lt = s.find('<', i);
            if (lt == std::string::npos) return fail_at("missing closing tag");
            text += s.substr(i, lt - i);
            i = lt;
            if (starts("</")) {
                i += 2;
                size_t gt = s.find('>', i);
                if (gt == std::string::npos) return fail_at("unterminated closing tag");
                if (trim(s.substr(i, gt - i)) != n.name) return fail_at(("mismatched closing tag for <" + n.name + ">").c_str());
                i = gt + 1;
                n.text = decode(trim(text));
                return true;
            }
            if (starts("<!--")) {
                if (!skip_past("-->")) return fail_at("unterminated comment");
                continue;
            }
            if (starts("<![CDATA[")) {
                i += 9;
                size_t e = s.find("]]>", i);
                if (e == std::string::npos) return fail_at("unterminated CDATA");
                text += s.substr(i, e - i);
                i = e + 3;
                continue;
            }
            if (starts("<?")) {
                if (!skip_past("?>")) return fail_at("unterminated processing instruction");
                continue;
            }
            n.kids.emplace_back();
            if (!element(n.kids.back(), depth + 1)) return false;
        }
    }
    bool document(XNode& root) {
        if (!skip_misc()) return false;
        if (!element(root, 0)) return false;
        if (!skip_misc()) return false;
        if (i != s.size()) return fail_at("content after the root element");
        return true;
    }
};

// CustomFileReader::get_float3 (customfilereader.cpp:342-362): three numbers
// separated by any of " {,}".
bool parse_float3(const std::string& t, float out[3]) {
    std::string u = t;
    for (char& c : u)
        if (c == '{' || c == '}' || c == ',') c = ' ';
    std::istringstream is(u);
    is.imbue(std::locale::classic());
    return bool(is >> out[0] >> out[1] >> out[2]);
}

// ------------------------------------------------------------------ OBJ --

struct ObjObject {
    std::string name;
    std::vector<uint32_t> tris;   // file-global 0-based vertex indices
    std::vector<uint32_t> ntris;  // file-global normal indices per corner (kNoIndex: none)
    std::vector<uint32_t> ttris;  // file-global texcoord indices per corner (kNoIndex: none)
};

struct ObjData {
    std::vector<float> verts, normals, texcoords;
    std::vector<ObjObject> objs;
};

// One OBJ index (1-based, or negative = relative to the end of the list so
// far) resolved against a list of n entries.
bool resolve_index(const char* tok, const char* end, size_t n, uint32_t* out) {
    char* e = nullptr;
    long v = std::strtol(tok, &e, 10);
    if (e == tok || e != end || v == 0) return false;
    long r = v > 0 ? v - 1 : (long)n + v;
    if (r < 0 || (size_t)r >= n) return false;
    *out = (uint32_t)r;
    return true;
}

// A face corner: v, v/vt, v//vn or v/vt/vn (ObjFileReader::parse_face,
// src/io/objfilereader.cpp:57-108, which reads v//vn and v/vt the other way
// round -- a slip not kept).  Missing fields give kNoIndex.
bool parse_corner(const std::string& c, const ObjData& D, uint32_t* v, uint32_t* vt, uint32_t* vn) {
    const size_t s1 = c.find('/');
    const size_t s2 = s1 == std::string::npos ? std::string::npos : c.find('/', s1 + 1);
    const char* b = c.c_str();
    const size_t vend = s1 == std::string::npos ? c.size() : s1;
    if (!resolve_index(b, b + vend, D.verts.size() / 3, v)) return false;
    *vt = *vn = kNoIndex;
    if (s1 == std::string::npos) return true;
    const size_t tend = s2 == std::string::npos ? c.size() : s2;
    if (tend > s1 + 1 && !resolve_index(b + s1 + 1, b + tend, D.texcoords.size() / 2, vt)) return false;
    if (s2 == std::string::npos) return tend > s1 + 1;  // "v/" is malformed
    return resolve_index(b + s2 + 1, b + c.size(), D.normals.size() / 3, vn);
}

rb_status read_obj(const std::string& path, ObjData& D) {
    std::vector<float>& verts = D.verts;
    std::vector<ObjObject>& objs = D.objs;
    std::ifstream in(path);
    if (!in) return fail(RB_ERR_INVALID, "cannot open %s", path.c_str());
    std::string line;
    size_t lineno = 0;
    objs.clear();
    auto current = [&]() -> ObjObject& {
        if (objs.empty()) objs.emplace_back();
        return objs.back();
    };
    std::vector<uint32_t> poly;
    while (std::getline(in, line)) {
        lineno++;
        size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream ls(line);
        ls.imbue(std::locale::classic());
        std::string tag;
        if (!(ls >> tag)) continue;
        if (tag == "o") {
            ObjObject o;
            ls >> o.name;
            objs.push_back(std::move(o));
        } else if (tag == "v" || tag == "vn") {
            float x, y, z;
            if (!(ls >> x >> y >> z))
                return fail(RB_ERR_INVALID, "%s:%zu: bad %s", path.c_str(), lineno, tag == "v" ? "vertex" : "normal");
            std::vector<float>& dst = tag == "v" ? verts : D.normals;
            dst.push_back(x);
            dst.push_back(y);
            dst.push_back(z);
        } else if (tag == "vt") {
            float u, v = 0.0f;  // a 1D coordinate gets v = 0; a third (w) is ignored
            if (!(ls >> u)) return fail(RB_ERR_INVALID, "%s:%zu: bad texture coordinate", path.c_str(), lineno);
            if (!(ls >> v)) v = 0.0f;
            D.texcoords.push_back(u);
            D.texcoords.push_back(v);
        } else if (tag == "f") {
            poly.clear();
            std::string c;
            while (ls >> c) {
                uint32_t vi, ti, ni;
                if (!parse_corner(c, D, &vi, &ti, &ni))
                    return fail(RB_ERR_INVALID, "%s:%zu: face corner '%s' is malformed or names no entry", path.c_str(),
                                lineno, c.c_str());
                poly.push_back(vi);
                poly.push_back(ti);
                poly.push_back(ni);
            }
            if (poly.size() < 9) return fail(RB_ERR_INVALID, "%s:%zu: face with %zu corners", path.c_str(), lineno,
                                             poly.size() / 3);
            ObjObject& o = current();
            const size_t nc = poly.size() / 3;
            // a triangle has normals (texcoords) only when all three corners do
            for (size_t k = 1; k + 1 < nc; k++) {
                const size_t cs[3] = {0, k, k + 1};
                bool has_t = true, has_n = true;
                for (size_t c3 : cs) {
                    o.tris.push_back(poly[3 * c3]);
                    has_t = has_t && poly[3 * c3 + 1] != kNoIndex;
                    has_n = has_n && poly[3 * c3 + 2] != kNoIndex;
                }
                for (size_t c3 : cs) {
                    o.ttris.push_back(has_t ? poly[3 * c3 + 1] : kNoIndex);
                    o.ntris.push_back(has_n ? poly[3 * c3 + 2] : kNoIndex);
                }
            }
        }
        // g, s, usemtl, mtllib: not used by the renderer
    }
    return RB_OK;
}

// Adds every object of a read file as a mesh of the file's vertices,
// normals and texture coordinates (file-global indexing, Scene::add_mesh,
// src/core/scene.cpp:83-113), object k with material obj_mat[k].
rb_status add_obj_to_scene(Scene& S, const ObjData& D, const std::vector<uint32_t>& obj_mat) {
    std::vector<uint32_t> idx, nidx, tidx, mats;
    size_t ntris = 0;
    for (const ObjObject& o : D.objs) ntris += o.tris.size() / 3;
    rb_status st = scene_room(S, D.verts.size() / 3, ntris);  // before anything is copied
    if (st != RB_OK) return st;
    try {
        idx.reserve(3 * ntris);
        nidx.reserve(3 * ntris);
        tidx.reserve(3 * ntris);
        mats.reserve(ntris);
        for (size_t k = 0; k < D.objs.size(); k++) {
            const ObjObject& o = D.objs[k];
            idx.insert(idx.end(), o.tris.begin(), o.tris.end());
            nidx.insert(nidx.end(), o.ntris.begin(), o.ntris.end());
            tidx.insert(tidx.end(), o.ttris.begin(), o.ttris.end());
            mats.insert(mats.end(), o.tris.size() / 3, obj_mat[k]);
        }
    } catch (...) {
        return fail(RB_ERR_OOM, "OBJ read failed (host allocation)");
    }
    const uint32_t nn = (uint32_t)(D.normals.size() / 3), ntc = (uint32_t)(D.texcoords.size() / 2);
    return append_mesh(S, D.verts.data(), (uint32_t)(D.verts.size() / 3), idx.data(), (uint32_t)ntris,
                       nn ? D.normals.data() : nullptr, nn, nn ? nidx.data() : nullptr,
                       ntc ? D.texcoords.data() : nullptr, ntc, ntc ? tidx.data() : nullptr, mats.data(), 0);
}

std::string dir_of(const std::string& p) {
    size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string() : p.substr(0, k + 1);
}

}  // namespace
}  // namespace rb

using namespace rb;

extern "C" {

rb_status rb_scene_load_obj(rb_scene_t s, const char* path, uint32_t material_id) {
    if (!s || !path) return fail(RB_ERR_INVALID, "NULL argument");
    if (material_id >= s->s.mats.size()) return fail(RB_ERR_INVALID, "material id out of range (add the material first)");
    ObjData D;
    rb_status st;
    try {
        st = read_obj(path, D);
    } catch (...) {
        return fail(RB_ERR_OOM, "OBJ read failed (host allocation)");
    }
    if (st != RB_OK) return st;
    return add_obj_to_scene(s->s, D, std::vector<uint32_t>(D.objs.size(), material_id));
}

rb_status rb_scene_load_xml(rb_scene_t s, const char* path, uint32_t width, uint32_t height, rb_camera* cam) {
    if (!s || !path) return fail(RB_ERR_INVALID, "NULL argument");
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(RB_ERR_INVALID, "cannot open %s", path);
    std::string src((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    XNode doc;
    XmlParser P(src);
    if (!P.document(doc)) return fail(RB_ERR_INVALID, "%s: %s", path, P.err.c_str());
    if (doc.name != "scene") return fail(RB_ERR_INVALID, "%s: root element is <%s>, not <scene>", path, doc.name.c_str());
    const XNode* files = doc.child("files");
    const XNode* file = files ? files->child("file") : nullptr;
    if (!file || file->text.empty()) return fail(RB_ERR_INVALID, "%s: no <files><file> mesh", path);
    const XNode* count = files->child("count");
    if (count && std::atoi(count->text.c_str()) != 1)
        return fail(RB_ERR_UNSUPPORTED, "%s: one mesh file per scene (count %s)", path, count->text.c_str());

    // materials in file order (customfilereader.cpp:89-225), then one emitter
    // per light (:254-280), then the layers of the layered materials
    Scene& S = s->s;
    const uint32_t mat_base = (uint32_t)S.mats.size();
    std::vector<rb_material> mats;
    std::map<std::string, uint32_t> by_name;
    // one <bxdf> (customfilereader.cpp:113-219)
    auto parse_bxdf = [&](const XNode* bx, const std::string& mname, rb_material* out) -> rb_status {
        const XNode* type = bx->child("type");
        const XNode* refl = bx->child("reflectance");
        rb_material mat{};
        if (!type) return fail(RB_ERR_INVALID, "%s: material '%s' bxdf without <type>", path, mname.c_str());
        auto num = [&](const char* tag, float* v) -> bool {
            const XNode* e = bx->child(tag);
            if (!e) return false;
            std::istringstream is(e->text);
            is.imbue(std::locale::classic());
            return bool(is >> *v);
        };
        bool ok = true;
        if (type->text == "lambertian") {
            mat.kind = RB_MAT_LAMBERT;
        } else if (type->text == "mirror") {
            mat.kind = RB_MAT_MIRROR;
        } else if (type->text == "glass") {  // customfilereader.cpp:160-174
            mat.kind = RB_MAT_GLASS;
            ok = num("etai", &mat.param[0]) && num("etat", &mat.param[1]);
        } else if (type->text == "orennayar") {  // :130-146 (A, B derived when the material is added)
            mat.kind = RB_MAT_ORENNAYAR;
            ok = num("sigma", &mat.param[0]);
        } else if (type->text == "microfacet") {  // :175-218
            mat.kind = RB_MAT_MICROFACET;
            const XNode* dist = bx->child("distribution");
            if (dist && dist->text != "blinn")
                return fail(RB_ERR_UNSUPPORTED, "%s: material '%s': microfacet distribution '%s' (blinn only)", path,
                            mname.c_str(), dist->text.c_str());
            ok = num("exponent", &mat.param[0]);
            const XNode* ft = bx->child("fresnel_type");
            const std::string f = ft ? ft->text : "fresnel_noop";
            if (f == "fresnel_conductor") {
                mat.param[1] = (float)RB_FRESNEL_CONDUCTOR;
                const XNode* eta = bx->child("eta");
                const XNode* k = bx->child("k");
                ok = ok && eta && k && parse_float3(eta->text, &mat.param[2]) && parse_float3(k->text, &mat.param[5]);
            } else if (f == "fresnel_dielectric") {
                mat.param[1] = (float)RB_FRESNEL_DIELECTRIC;
                ok = ok && num("etai", &mat.param[2]) && num("etat", &mat.param[3]);
            } else if (f == "fresnel_noop") {
                mat.param[1] = (float)RB_FRESNEL_NONE;
            } else {
                return fail(RB_ERR_INVALID, "%s: material '%s': unknown fresnel_type '%s'", path, mname.c_str(),
                            f.c_str());
            }
        } else {
            return fail(RB_ERR_UNSUPPORTED, "%s: material '%s': bxdf type '%s' is not supported", path, mname.c_str(),
                        type->text.c_str());
        }
        if (!ok) return fail(RB_ERR_INVALID, "%s: material '%s': missing or bad %s parameters", path, mname.c_str(),
                             type->text.c_str());
        if (!refl || !parse_float3(refl->text, mat.color))
            return fail(RB_ERR_INVALID, "%s: material '%s' needs a <reflectance> triple", path, mname.c_str());
        *out = mat;
        return RB_OK;
    };
    std::vector<std::vector<rb_material>> layered;  // the layers of each named material with more than one
    for (const XNode& m : doc.kids) {
        if (m.name != "material") continue;
        const XNode* name = m.child("name");
        const XNode* bsdf = m.child("bsdf");
        if (!name || !bsdf) return fail(RB_ERR_INVALID, "%s: <material> needs <name> and <bsdf>", path);
        std::vector<rb_material> layers;
        for (const XNode& k : bsdf->kids) {
            if (k.name != "bxdf") continue;
            rb_material l;
            const rb_status st = parse_bxdf(&k, name->text, &l);
            if (st != RB_OK) return st;
            layers.push_back(l);
        }
        // <layers> gives the count (material.num_valid_layers, :104-110)
        const XNode* nl = bsdf->child("layers");
        if (nl && std::atoi(nl->text.c_str()) != (int)layers.size())
            return fail(RB_ERR_INVALID, "%s: material '%s': <layers> says %s, %zu <bxdf> given", path,
                        name->text.c_str(), nl->text.c_str(), layers.size());
        if (layers.empty() || layers.size() > 4)
            return fail(RB_ERR_INVALID, "%s: material '%s' has %zu layers (1..4)", path, name->text.c_str(),
                        layers.size());
        by_name[name->text] = mat_base + (uint32_t)mats.size();
        if (layers.size() == 1) {
            mats.push_back(layers[0]);
        } else {
            rb_material lm{};
            lm.kind = RB_MAT_LAYERED;
            lm.param[0] = (float)layers.size();
            mats.push_back(lm);  // layer ids set below, once the lights are numbered
            layered.push_back(layers);
        }
    }
    std::map<std::string, uint32_t> light_mat;
    if (const XNode* lo = doc.child("lightobjects")) {
        for (const XNode& l : lo->kids) {
            if (l.name != "light") continue;
            const XNode* name = l.child("name");
            const XNode* le = l.child("Le");
            rb_material mat{};
            mat.kind = RB_MAT_EMITTER;
            if (!name || !le || !parse_float3(le->text, mat.color))
                return fail(RB_ERR_INVALID, "%s: <light> needs <name> and an <Le> triple", path);
            light_mat[name->text] = mat_base + (uint32_t)mats.size();
            mats.push_back(mat);
        }
    }
    if (mats.empty()) return fail(RB_ERR_INVALID, "%s: no materials", path);
    for (size_t i = 0, li = 0; i < mats.size(); i++) {
        if (mats[i].kind != RB_MAT_LAYERED) continue;
        const std::vector<rb_material>& layers = layered[li++];
        for (size_t k = 0; k < layers.size(); k++) {
            mats[i].param[1 + k] = (float)(mat_base + mats.size());
            mats.push_back(layers[k]);
        }
    }

    ObjData D;
    std::vector<ObjObject>& objs = D.objs;
    rb_status st;
    try {
        st = read_obj(dir_of(path) + file->text, D);
    } catch (...) {
        return fail(RB_ERR_OOM, "OBJ read failed (host allocation)");
    }
    if (st != RB_OK) return st;
    // object -> material (customfilereader.cpp:227-246); unassigned objects
    // keep material 0 of this file, lights override (:282-298)
    std::vector<uint32_t> obj_mat(objs.size(), mat_base);
    for (const XNode& o : doc.kids) {
        if (o.name != "object") continue;
        const XNode* name = o.child("name");
        const XNode* mat = o.child("material");
        if (!name || !mat) return fail(RB_ERR_INVALID, "%s: <object> needs <name> and <material>", path);
        auto it = by_name.find(mat->text);
        if (it == by_name.end())
            return fail(RB_ERR_INVALID, "%s: object '%s' names unknown material '%s'", path, name->text.c_str(),
                        mat->text.c_str());
        for (size_t k = 0; k < objs.size(); k++)
            if (objs[k].name == name->text) {
                obj_mat[k] = it->second;
                break;
            }
    }
    for (size_t k = 0; k < objs.size(); k++) {
        auto it = light_mat.find(objs[k].name);
        if (it != light_mat.end()) obj_mat[k] = it->second;
    }
    rb_camera c{};
    if (cam) {
        const XNode* e = doc.child("camera");
        const XNode* pos = e ? e->child("position") : nullptr;
        const XNode* at = e ? e->child("lookat") : nullptr;
        const XNode* up = e ? e->child("up") : nullptr;
        if (!pos || !at || !up || !parse_float3(pos->text, c.position) || !parse_float3(at->text, c.lookat) ||
            !parse_float3(up->text, c.up))
            return fail(RB_ERR_INVALID, "%s: <camera> needs <position>, <lookat> and <up> triples", path);
        const XNode* fov = e->child("fov");
        c.hfov_deg = fov && !fov->text.empty() ? (float)std::atof(fov->text.c_str()) : 90.0f;
        c.width = width;
        c.height = height;
    }
    // validation and derived parameters as for any material: the single
    // BSDFs first, then the layered ones against the full list
    for (rb_material& m : mats)
        if (m.kind != RB_MAT_LAYERED && (st = prepare_material(S.mats, &m)) != RB_OK) return st;
    {
        std::vector<rb_material> all(S.mats);
        all.insert(all.end(), mats.begin(), mats.end());
        for (rb_material& m : mats)
            if (m.kind == RB_MAT_LAYERED && (st = prepare_material(all, &m)) != RB_OK) return st;
    }
    S.mats.insert(S.mats.end(), mats.begin(), mats.end());
    if ((st = add_obj_to_scene(S, D, obj_mat)) != RB_OK) {
        S.mats.resize(mat_base);
        return st;
    }
    if (cam) *cam = c;
    return RB_OK;
}

}  // extern "C"
