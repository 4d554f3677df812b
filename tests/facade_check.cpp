/*
 * facade_check.cpp - exercises the C++ API (include/exposurerender.h) the way an application
 * written against the reference would, and cross-checks it against the C-ABI.
 *
 *   er_facade_check nodevice   without a GPU: a bind must throw Exception(Error)
 *   er_facade_check render     with a GPU: one scene bound through the C++ API and the same
 *                              scene bound through the C-ABI render identical images; ID,
 *                              rebind, unbind, capacity and exception semantics
 * Prints one line per check and exits non-zero on the first failure.
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "exposurerender.h"

using namespace ExposureRender;

static int g_fail = 0;
#define CHECK(cond, ...)                                        \
    do {                                                        \
        if (cond) { std::printf("ok   "); }                     \
        else { std::printf("FAIL "); g_fail = 1; }              \
        std::printf(__VA_ARGS__);                               \
        std::printf("\n");                                      \
        std::fflush(stdout);                                    \
    } while (0)

static const int R = 32, FILM = 64, PASSES = 4;

/* a plane light of side 0.5 at (0, 0, 1) facing -z: TM = translate(0,0,1) * rot_x(pi) */
static void plane_light_down(float m[4][4], float inv[4][4])
{
    const float tm[4][4] = {{1, 0, 0, 0}, {0, -1, 0, 0}, {0, 0, -1, 1}, {0, 0, 0, 1}};
    const float it[4][4] = {{1, 0, 0, 0}, {0, -1, 0, 0}, {0, 0, -1, 1}, {0, 0, 0, 1}};
    std::memcpy(m, tm, sizeof(tm));
    std::memcpy(inv, it, sizeof(it));
}

static int nodevice()
{
    std::vector<unsigned short> vox(8 * 8 * 8, 1000);
    ErVolume V;
    V.BindVoxels(Vec3i(8, 8, 8), Vec3f(1.0f), vox.data(), true);
    try {
        BindVolume(V);
        CHECK(false, "BindVolume without a device must throw");
    } catch (const Exception& e) {
        CHECK(e.Level == Enums::Error, "BindVolume without a device throws level %d: %s", (int)e.Level, e.Message);
    }
    return g_fail;
}

static int render()
{
    std::vector<unsigned short> vox((size_t)R * R * R);
    if (er_phantom_blobs(vox.data(), R, R, R, 3, 0.1f, 0.25f, 1000.0f, 4000.0f, 50.0f, 11) != ER_OK) return 2;

    /* ---- scene through the C++ API ---- */
    ErVolume V;
    V.BindVoxels(Vec3i(R, R, R), Vec3f(1.0f), vox.data(), true);
    BindVolume(V);
    CHECK(V.ID >= 0, "BindVolume assigns an ID (%d)", V.ID);
    CHECK(!V.Voxels.Dirty, "BindVolume clears the Dirty flag");

    ErTexture Tex;
    Tex.Procedural.UniformColor = ColorXYZf(1.0f, 1.0f, 1.0f);
    BindTexture(Tex);

    ErLight L;
    L.Visible = true;
    L.TextureID = Tex.ID;
    L.Multiplier = 50.0f;
    L.Shape.Type = Enums::Plane;
    L.Shape.Size = Vec3f(0.5f, 0.5f, 0.0f);
    plane_light_down(L.Shape.TM.NN, L.Shape.InvTM.NN);
    BindLight(L);

    ErTracer T;
    T.Opacity1D.AddNode(ScalarNode(0.0f, 0.0f));
    T.Opacity1D.AddNode(ScalarNode(400.0f, 0.0f));
    T.Opacity1D.AddNode(ScalarNode(1500.0f, 0.6f));
    T.Opacity1D.AddNode(ScalarNode(4000.0f, 1.0f));
    T.Diffuse1D.AddNode(ColorNode(0.0f, 0.9f, 0.6f, 0.5f));
    T.Diffuse1D.AddNode(ColorNode(4000.0f, 1.0f, 1.0f, 1.0f));
    T.Specular1D.AddNode(ColorNode(0.0f, 0.1f, 0.1f, 0.1f));
    T.Specular1D.AddNode(ColorNode(4000.0f, 0.3f, 0.3f, 0.3f));
    T.Glossiness1D.AddNode(ScalarNode(0.0f, 0.5f));
    T.Glossiness1D.AddNode(ScalarNode(4000.0f, 0.5f));
    T.Emission1D.AddNode(ColorNode(0.0f, 0.0f, 0.0f, 0.0f));
    T.Camera.FilmSize = Vec2i(FILM, FILM);
    T.Camera.Pos = Vec3f(1.2f, 0.9f, 1.6f);
    T.Camera.Target = Vec3f(0.0f, 0.0f, 0.0f);
    T.Camera.Up = Vec3f(0.0f, 1.0f, 0.0f);
    T.Camera.FocalDistance = -1.0f;
    T.Camera.ApertureSize = 0.0f;
    T.Camera.ClipNear = 0.0f;
    T.Camera.ClipFar = 1000.0f;
    T.Camera.Exposure = 4.0f;
    T.Camera.Gamma = 2.2f;
    T.Camera.FOV = 35.0f;
    T.RenderSettings.Shading.DensityScale = 10.0f;
    T.VolumeID = V.ID;
    T.LightIDs[0] = L.ID;
    T.LightIDs.Count = 1;
    BindTracer(T);
    CHECK(T.ID >= 0, "BindTracer assigns an ID (%d)", T.ID);

    /* ---- the same scene through the C-ABI ---- */
    er_tracer_desc d;
    er_tracer_desc_init(&d);
    er_plf_add_node(&d.opacity, 0.0f, 0.0f);
    er_plf_add_node(&d.opacity, 400.0f, 0.0f);
    er_plf_add_node(&d.opacity, 1500.0f, 0.6f);
    er_plf_add_node(&d.opacity, 4000.0f, 1.0f);
    const float dif0[3] = {0.9f, 0.6f, 0.5f}, spc0[3] = {0.1f, 0.1f, 0.1f};
    for (int c = 0; c < 3; ++c) {
        er_plf_add_node(&d.diffuse.plf[c], 0.0f, dif0[c]);
        er_plf_add_node(&d.diffuse.plf[c], 4000.0f, 1.0f);
        er_plf_add_node(&d.specular.plf[c], 0.0f, spc0[c]);
        er_plf_add_node(&d.specular.plf[c], 4000.0f, 0.3f);
        er_plf_add_node(&d.emission.plf[c], 0.0f, 0.0f);
    }
    er_plf_add_node(&d.glossiness, 0.0f, 0.5f);
    er_plf_add_node(&d.glossiness, 4000.0f, 0.5f);
    d.camera.film_size[0] = d.camera.film_size[1] = FILM;
    d.camera.pos[0] = 1.2f; d.camera.pos[1] = 0.9f; d.camera.pos[2] = 1.6f;
    d.camera.up[1] = 1.0f;
    d.camera.focal_distance = -1.0f;
    d.camera.clip_far = 1000.0f;
    d.camera.exposure = 4.0f;
    d.camera.gamma = 2.2f;
    d.camera.fov = 35.0f;
    d.settings.density_scale = 10.0f;
    d.volume_id = V.ID;
    d.light_ids[0] = L.ID;
    d.light_count = 1;
    CHECK(er_bind_tracer(&d, 1) == ER_OK && d.id >= 0 && d.id != T.ID, "C-ABI tracer bound (%d)", d.id);

    er_seed_tracer(T.ID, 77);
    er_seed_tracer(d.id, 77);
    for (int k = 0; k < PASSES; ++k) RenderEstimate(T.ID);
    for (int k = 0; k < PASSES; ++k) er_render_estimate(d.id);
    std::vector<unsigned char> a((size_t)FILM * FILM * 4), b(a.size());
    GetEstimate(T.ID, a.data());
    er_get_estimate(d.id, b.data(), b.size());
    size_t nz = 0;
    for (unsigned char x : a) nz += x != 0;
    CHECK(a == b, "C++ API and C-ABI images are identical (%zu non-zero bytes)", nz);
    CHECK(nz > 0, "the image is not empty");

    int n = -1;
    GetNoIterations(T.ID, n);
    CHECK(n == PASSES, "GetNoIterations = %d", n);

    /* ---- a bitmap-textured floor object, full estimator (ObjectIDs name the objects) ---- */
    const int BW = 8, BH = 4;
    std::vector<ColorRGBAuc> pix((size_t)BW * BH);
    for (size_t i = 0; i < pix.size(); ++i)
        for (int c = 0; c < 4; ++c) pix[i].D[c] = (unsigned char)((i * 37 + c * 91) & 255);
    /* a 3 x 3 plane at y = -0.6 facing +y: local x -> x, y -> -z, z -> y */
    const float ftm[4][4] = {{1, 0, 0, 0}, {0, 0, 1, -0.6f}, {0, -1, 0, 0}, {0, 0, 0, 1}};
    const float finv[4][4] = {{1, 0, 0, 0}, {0, 0, -1, 0}, {0, 1, 0, 0.6f}, {0, 0, 0, 1}};

    ErBitmap B;
    B.BindPixels(Vec2i(BW, BH), pix.data());
    BindBitmap(B);
    CHECK(B.ID >= 0 && !B.Pixels.Dirty, "BindBitmap assigns an ID (%d) and clears Dirty", B.ID);
    ErTexture BT;
    BT.Type = Enums::Bitmap;
    BT.BitmapID = B.ID;
    BT.Repeat = Vec2f(2.0f, 1.0f);
    BindTexture(BT);
    ErObject Fl;
    Fl.DiffuseTextureID = BT.ID;
    Fl.SpecularTextureID = Tex.ID;
    Fl.Shape.Type = Enums::Plane;
    Fl.Shape.Size = Vec3f(3.0f, 3.0f, 0.0f);
    std::memcpy(Fl.Shape.TM.NN, ftm, sizeof(ftm));
    std::memcpy(Fl.Shape.InvTM.NN, finv, sizeof(finv));
    BindObject(Fl);
    CHECK(Fl.ID >= 0, "BindObject assigns an ID (%d)", Fl.ID);
    T.RenderSettings.Estimator = Enums::Full;
    T.ObjectIDs[0] = Fl.ID;
    T.ObjectIDs.Count = 1;
    BindTracer(T);

    er_bitmap_desc bd;
    er_bitmap_desc_init(&bd);
    bd.resolution[0] = BW; bd.resolution[1] = BH;
    bd.pixels = reinterpret_cast<const uint8_t*>(pix.data());
    bd.pixels_dirty = 1;
    CHECK(er_bind_bitmap(&bd, 1) == ER_OK && bd.id >= 0 && bd.id != B.ID, "C-ABI bitmap bound (%d)", bd.id);
    er_texture_desc td;
    er_texture_desc_init(&td);
    td.type = ER_TEXTURE_BITMAP;
    td.bitmap_id = bd.id;
    td.repeat[0] = 2.0f; td.repeat[1] = 1.0f;
    CHECK(er_bind_texture(&td, 1) == ER_OK, "C-ABI bitmap texture bound (%d)", td.id);
    er_object_desc od;
    er_object_desc_init(&od);
    od.diffuse_texture_id = td.id;
    od.specular_texture_id = Tex.ID;
    od.shape.type = ER_SHAPE_PLANE;
    od.shape.size[0] = od.shape.size[1] = 3.0f;
    std::memcpy(od.shape.tm, ftm, sizeof(ftm));
    std::memcpy(od.shape.inv_tm, finv, sizeof(finv));
    CHECK(er_bind_object(&od, 1) == ER_OK && od.id >= 0, "C-ABI object bound (%d)", od.id);
    d.settings.estimator = ER_ESTIMATOR_FULL;
    d.object_ids[0] = od.id;
    d.object_count = 1;
    CHECK(er_bind_tracer(&d, 1) == ER_OK, "C-ABI tracer rebound with the object");

    er_seed_tracer(T.ID, 78);
    er_seed_tracer(d.id, 78);
    for (int k = 0; k < PASSES; ++k) RenderEstimate(T.ID);
    for (int k = 0; k < PASSES; ++k) er_render_estimate(d.id);
    GetEstimate(T.ID, a.data());
    er_get_estimate(d.id, b.data(), b.size());
    size_t nz2 = 0;
    for (unsigned char x : a) nz2 += x != 0;
    CHECK(a == b, "objects and bitmaps: C++ API and C-ABI images are identical (%zu non-zero bytes)", nz2);
    CHECK(nz2 > nz, "the floor object adds coverage (%zu > %zu)", nz2, nz);

    /* rebind with the same ID updates in place and restarts accumulation */
    const int id0 = T.ID;
    T.Camera.Exposure = 2.0f;
    BindTracer(T);
    GetNoIterations(T.ID, n);
    CHECK(T.ID == id0 && n == 0, "rebind keeps ID %d and resets NoIterations (%d)", T.ID, n);

    /* unknown IDs raise Warning, as the reference's list lookups */
    try {
        RenderEstimate(9999);
        CHECK(false, "RenderEstimate(9999) must throw");
    } catch (const Exception& e) {
        CHECK(e.Level == Enums::Warning, "RenderEstimate(unknown) throws Warning: %s", e.Message);
    }

    /* a 32nd live light is ignored and keeps ID -1 (31 live objects per kind) */
    std::vector<ErLight> many(40, L);
    int assigned = 1, ignored = 0;   /* L is already live */
    for (auto& x : many) {
        x.ID = -1;
        BindLight(x);
        if (x.ID >= 0) ++assigned; else ++ignored;
    }
    CHECK(assigned == 31 && ignored == 10, "light capacity: %d live, %d ignored", assigned, ignored);
    for (auto& x : many) if (x.ID >= 0) BindLight(x, false);

    /* unbind; a second unbind of the same ID is ignored */
    BindTracer(T, false);
    BindTracer(T, false);
    try {
        RenderEstimate(id0);
        CHECK(false, "RenderEstimate after unbind must throw");
    } catch (const Exception& e) {
        CHECK(e.Level == Enums::Warning, "RenderEstimate after unbind throws Warning");
    }
    er_shutdown();
    return g_fail;
}

int main(int argc, char** argv)
{
    if (argc < 2) { std::fprintf(stderr, "usage: %s nodevice|render\n", argv[0]); return 2; }
    try {
        if (!std::strcmp(argv[1], "nodevice")) return nodevice();
        if (!std::strcmp(argv[1], "render")) return render();
    } catch (const Exception& e) {
        std::printf("FAIL uncaught Exception level %d: %s\n", (int)e.Level, e.Message);
        return 1;
    }
    return 2;
}
