// rb_camera.cpp -- PerspectiveCamera's transform chain on the host
// (renderbox2 src/core/camera.cu:30-57, src/core/transform.cu:28-126).
// Matrix products are row-by-column float4 dot products summed left to
// right; inverses use the cofactor expansion of Matrix4x4::invert.
#include "rb_internal.h"

namespace rb {

void invert4x4_ref(const float* m, float* out);

namespace {

struct M4 {
    float m[16];
};
struct Xf {
    M4 m, inv;
};

M4 ident() {
    M4 r{};
    for (int i = 0; i < 4; i++) r.m[5 * i] = 1.0f;
    return r;
}
M4 mul(const M4& a, const M4& b) {
    M4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            r.m[4 * i + j] = a.m[4 * i] * b.m[j] + a.m[4 * i + 1] * b.m[4 + j] + a.m[4 * i + 2] * b.m[8 + j] +
                             a.m[4 * i + 3] * b.m[12 + j];
    return r;
}
M4 inv(const M4& a) {
    M4 r;
    invert4x4_ref(a.m, r.m);
    return r;
}
Xf xf(const M4& m) { return Xf{m, inv(m)}; }
Xf xmul(const Xf& a, const Xf& b) { return Xf{mul(a.m, b.m), mul(b.inv, a.inv)}; }
Xf xinv(const Xf& a) { return Xf{a.inv, a.m}; }

Xf scale(float x, float y, float z) {
    M4 r = ident();
    r.m[0] = x;
    r.m[5] = y;
    r.m[10] = z;
    r.m[15] = 1.0f;
    return xf(r);
}
Xf translate(float x, float y, float z) {
    M4 r = ident();
    r.m[3] = x;
    r.m[7] = y;
    r.m[11] = z;
    return xf(r);
}
// lookat (transform.cu:78-107)
Xf lookat(F3 pos, F3 at, F3 up) {
    F3 dir = normalize(at - pos);
    F3 left = normalize(cross(normalize(up), dir));
    F3 nup = cross(dir, left);
    M4 c2w;
    const float m[16] = {left.x, nup.x, dir.x, pos.x, left.y, nup.y, dir.y, pos.y,
                         left.z, nup.z, dir.z, pos.z, 0.0f,   0.0f,  0.0f,  1.0f};
    std::memcpy(c2w.m, m, sizeof(m));
    return Xf{inv(c2w), c2w};
}
// perspective (transform.cu:114-124)
Xf perspective(float fov, float n, float f) {
    M4 p{};
    p.m[0] = 1.0f;
    p.m[5] = 1.0f;
    p.m[10] = f / (f - n);
    p.m[11] = -f * n / (f - n);
    p.m[14] = 1.0f;
    const float to_rad = 3.14159265358979323846f / 180.f;
    float inv_tan = 1.0f / std::tan(to_rad * fov / 2.0f);
    return xmul(scale(inv_tan, inv_tan, 1.0f), xf(p));
}

}  // namespace
}  // namespace rb

using namespace rb;

extern "C" rb_status rb_camera_matrices(const rb_camera* c, float r2c[16], float c2w[16]) {
    if (!c || !r2c || !c2w) return fail(RB_ERR_INVALID, "NULL argument");
    if (c->width == 0 || c->height == 0) return fail(RB_ERR_INVALID, "empty film");
    if (!(c->hfov_deg > 0.0f && c->hfov_deg < 180.0f)) return fail(RB_ERR_INVALID, "hfov must be in (0, 180)");
    F3 pos = f3(c->position[0], c->position[1], c->position[2]);
    F3 at = f3(c->lookat[0], c->lookat[1], c->lookat[2]);
    F3 up = f3(c->up[0], c->up[1], c->up[2]);
    // The handedness flip is applied in camera space (scale * lookat).  The
    // reference multiplies lookat * scale (camera.cu:39), which mirrors the
    // world instead and moves a camera at x != 0 to -x; for cameras with
    // position.x == lookat.x == 0 and up = +y both orders give the same rays.
    Xf w2c = xmul(scale(-1.0f, 1.0f, 1.0f), lookat(pos, at, up));
    Xf cam2world = xinv(w2c);
    Xf cam2screen = perspective(c->hfov_deg, 1.0f, 1000.0f);
    Xf screen2raster = xmul(xmul(scale((float)c->width, (float)c->height, 1.0f), scale(0.5f, -0.5f, 1.0f)),
                            translate(1.0f, -1.0f, 0.0f));
    Xf raster2screen = xinv(screen2raster);
    Xf raster2cam = xmul(xinv(cam2screen), raster2screen);
    std::memcpy(r2c, raster2cam.m.m, 64);
    std::memcpy(c2w, cam2world.m.m, 64);
    return RB_OK;
}
