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
// reference numbers them.  Single-layer lambertian, mirror, glass,
// orennayar and (Blinn) microfacet bxdfs map to the renderer's materials;
// layered materials are refused with RB_ERR_UNSUPPORTED.  The camera's <fov> is used when present
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
    std::vector<uint32_t> tris;  // file-global 0-based vertex indices
};

bool resolve_index(const char* tok, size_t nv, uint32_t* out) {
    char* end = nullptr;
    long v = std::strtol(tok, &end, 10);
    if (end == tok || v == 0) return false;
    long r = v > 0 ? v - 1 : (long)nv + v;
    if (r < 0 || (size_t)r >= nv) return false;
    *out = (uint32_t)r;
    return true;
}

rb_status read_obj(const std::string& path, std::vector<float>& verts, std::vector<ObjObject>& objs) {
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
        } else if (tag == "v") {
            float x, y, z;
            if (!(ls >> x >> y >> z)) return fail(RB_ERR_INVALID, "%s:%zu: bad vertex", path.c_str(), lineno);
            verts.push_back(x);
            verts.push_back(y);
            verts.push_back(z);
        } else if (tag == "f") {
            poly.clear();
            std::string c;
            const size_t nv = verts.size() / 3;
            while (ls >> c) {
                // v, v/vt, v//vn, v/vt/vn: the vertex index is the first field
                uint32_t vi;
                if (!resolve_index(c.c_str(), nv, &vi))
                    return fail(RB_ERR_INVALID, "%s:%zu: face corner '%s' names no vertex", path.c_str(), lineno,
                                c.c_str());
                poly.push_back(vi);
            }
            if (poly.size() < 3) return fail(RB_ERR_INVALID, "%s:%zu: face with %zu corners", path.c_str(), lineno,
                                             poly.size());
            ObjObject& o = current();
            for (size_t k = 1; k + 1 < poly.size(); k++) {
                o.tris.push_back(poly[0]);
                o.tris.push_back(poly[k]);
                o.tris.push_back(poly[k + 1]);
            }
        }
        // vn, vt, g, s, usemtl, mtllib: not used by the renderer
    }
    return RB_OK;
}

rb_status add_obj_to_scene(Scene& S, const std::vector<float>& verts, const std::vector<ObjObject>& objs,
                           const std::vector<uint32_t>& obj_mat) {
    const uint32_t base = S.num_verts();
    size_t ntris = 0;
    for (const ObjObject& o : objs) ntris += o.tris.size() / 3;
    rb_status st = scene_room(S, verts.size() / 3, ntris);
    if (st != RB_OK) return st;
    try {
        S.verts.insert(S.verts.end(), verts.begin(), verts.end());
        for (size_t k = 0; k < objs.size(); k++) {
            const uint32_t tbase = S.num_tris();
            const uint32_t n = (uint32_t)(objs[k].tris.size() / 3);
            for (uint32_t v : objs[k].tris) S.idx.push_back(base + v);
            S.tri_mat.insert(S.tri_mat.end(), n, obj_mat[k]);
            if (S.mats[obj_mat[k]].kind == RB_MAT_EMITTER)
                for (uint32_t t = 0; t < n; t++) S.light_tris.push_back(tbase + t);
        }
    } catch (...) {
        return fail(RB_ERR_OOM, "scene growth failed");
    }
    return RB_OK;
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
    std::vector<float> verts;
    std::vector<ObjObject> objs;
    rb_status st;
    try {
        st = read_obj(path, verts, objs);
    } catch (...) {
        return fail(RB_ERR_OOM, "OBJ read failed (host allocation)");
    }
    if (st != RB_OK) return st;
    return add_obj_to_scene(s->s, verts, objs, std::vector<uint32_t>(objs.size(), material_id));
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

    // materials in file order (customfilereader.cpp:89-225), then one emitter per light (:254-280)
    Scene& S = s->s;
    const uint32_t mat_base = (uint32_t)S.mats.size();
    std::vector<rb_material> mats;
    std::map<std::string, uint32_t> by_name;
    for (const XNode& m : doc.kids) {
        if (m.name != "material") continue;
        const XNode* name = m.child("name");
        const XNode* bsdf = m.child("bsdf");
        if (!name || !bsdf) return fail(RB_ERR_INVALID, "%s: <material> needs <name> and <bsdf>", path);
        std::vector<const XNode*> bxdfs;
        for (const XNode& k : bsdf->kids)
            if (k.name == "bxdf") bxdfs.push_back(&k);
        if (bxdfs.size() != 1)
            return fail(RB_ERR_UNSUPPORTED, "%s: material '%s' has %zu layers; only single-layer materials are supported",
                        path, name->text.c_str(), bxdfs.size());
        const XNode* bx = bxdfs[0];
        const XNode* type = bx->child("type");
        const XNode* refl = bx->child("reflectance");
        rb_material mat{};
        if (!type) return fail(RB_ERR_INVALID, "%s: material '%s' bxdf without <type>", path, name->text.c_str());
        auto num = [&](const char* tag, float* out) -> bool {
            const XNode* e = bx->child(tag);
            if (!e) return false;
            std::istringstream is(e->text);
            is.imbue(std::locale::classic());
            return bool(is >> *out);
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
                            name->text.c_str(), dist->text.c_str());
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
                return fail(RB_ERR_INVALID, "%s: material '%s': unknown fresnel_type '%s'", path, name->text.c_str(),
                            f.c_str());
            }
        } else {
            return fail(RB_ERR_UNSUPPORTED, "%s: material '%s': bxdf type '%s' is not supported", path,
                        name->text.c_str(), type->text.c_str());
        }
        if (!ok) return fail(RB_ERR_INVALID, "%s: material '%s': missing or bad %s parameters", path,
                             name->text.c_str(), type->text.c_str());
        if (!refl || !parse_float3(refl->text, mat.color))
            return fail(RB_ERR_INVALID, "%s: material '%s' needs a <reflectance> triple", path, name->text.c_str());
        by_name[name->text] = mat_base + (uint32_t)mats.size();
        mats.push_back(mat);
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

    std::vector<float> verts;
    std::vector<ObjObject> objs;
    rb_status st;
    try {
        st = read_obj(dir_of(path) + file->text, verts, objs);
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
    for (const rb_material& m : mats) {  // validation and derived parameters as for any material
        if ((st = rb_scene_add_material(s, &m, nullptr)) != RB_OK) {
            S.mats.resize(mat_base);
            return st;
        }
    }
    if ((st = add_obj_to_scene(S, verts, objs, obj_mat)) != RB_OK) {
        S.mats.resize(mat_base);
        return st;
    }
    if (cam) *cam = c;
    return RB_OK;
}

}  // extern "C"
