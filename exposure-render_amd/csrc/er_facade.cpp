/*
 * er_facade.cpp - the reference's C++ API (Source/exposurerender.h:32-42) over the C-ABI.
 *
 * Each call converts the descriptor object into its plain-data mirror, calls the C-ABI entry
 * point and turns a non-OK status into ExposureRender::Exception with the level the status
 * encodes (status = ExceptionLevel + 1) and the message of er_last_error().
 */
#include "../../include/exposurerender.h"

#include <cstdint>

namespace ExposureRender
{
namespace
{

void check(int status)
{
    if (status == ER_OK) return;
    char msg[MAX_CHAR_SIZE];
    er_last_error(msg, sizeof(msg));
    int level = status - 1;
    if (level < Enums::Info || level > Enums::Fatal) level = Enums::Error;
    throw Exception((Enums::ExceptionLevel)level, msg);
}

void to_plf(const PiecewiseLinearFunction<MAX_NO_TF_NODES>& s, er_plf& d)
{
    d.node_range[0] = s.NodeRange[0];
    d.node_range[1] = s.NodeRange[1];
    for (int i = 0; i < ER_MAX_NO_TF_NODES; ++i) {
        d.position[i] = s.Position[i];
        d.value[i] = s.Value[i];
    }
    d.count = s.Count;
}

void to_ctf(const ColorTransferFunction1D& s, er_color_tf& d)
{
    for (int c = 0; c < 3; ++c) to_plf(s.PLF[c], d.plf[c]);
}

void to_shape(const Shape& s, er_shape& d)
{
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            d.tm[r][c] = s.TM.NN[r][c];
            d.inv_tm[r][c] = s.InvTM.NN[r][c];
        }
    d.one_sided = s.OneSided ? 1 : 0;
    d.type = (int)s.Type;
    for (int i = 0; i < 3; ++i) d.size[i] = s.Size[i];
    d.area = s.Area;
    d.inner_radius = s.InnerRadius;
    d.outer_radius = s.OuterRadius;
}

void to_indices(const Indices& s, int* ids, int& count)
{
    count = s.Count < 0 ? 0 : (s.Count > ER_MAX_INDICES ? ER_MAX_INDICES : s.Count);
    for (int i = 0; i < ER_MAX_INDICES; ++i) ids[i] = s.IDs[i];
}

}  // namespace

Camera::Camera()
{
    er_tracer_desc d;
    er_tracer_desc_init(&d);
    const er_camera& c = d.camera;
    FilmSize = Vec2i(c.film_size[0], c.film_size[1]);
    Pos = Vec3f(c.pos[0], c.pos[1], c.pos[2]);
    Target = Vec3f(c.target[0], c.target[1], c.target[2]);
    Up = Vec3f(c.up[0], c.up[1], c.up[2]);
    FocalDistance = c.focal_distance;
    ApertureSize = c.aperture_size;
    ClipNear = c.clip_near;
    ClipFar = c.clip_far;
    Exposure = c.exposure;
    Gamma = c.gamma;
    FOV = c.fov;
}

RenderSettings::RenderSettings()
{
    er_tracer_desc d;
    er_tracer_desc_init(&d);
    const er_render_settings& s = d.settings;
    Traversal.StepFactorPrimary = s.step_factor_primary;
    Traversal.StepFactorShadow = s.step_factor_shadow;
    Traversal.Shadows = s.shadows != 0;
    Traversal.MaxShadowDistance = s.max_shadow_distance;
    Shading.Type = s.shading_type;
    Shading.DensityScale = s.density_scale;
    Shading.OpacityModulated = s.opacity_modulated != 0;
    Shading.GradientComputation = s.gradient_computation;
    Shading.GradientThreshold = s.gradient_threshold;
    Shading.GradientFactor = s.gradient_factor;
    Estimator = (Enums::EstimatorMode)s.estimator;
    Sampler = (Enums::FreeFlightSampler)s.sampler;
    Bilateral = s.bilateral != 0;
    BilateralRadius = s.bilateral_radius;
    BilateralSigmaSpatial = s.bilateral_sigma_spatial;
    BilateralSigmaRange = s.bilateral_sigma_range;
}

Shape::Shape() : OneSided(false), Type(Enums::Plane), Area(0.0f), InnerRadius(0.0f), OuterRadius(0.0f) {}

/* Shape::Update (Source/shape.h:91-103): the runtime derives Area the same way at bind time */
void Shape::Update()
{
    const float PI = 3.141592654f;
    switch (Type) {
        case Enums::Plane:    Area = Size[0] * Size[1]; break;
        case Enums::Disk:     Area = PI * (OuterRadius * OuterRadius); break;
        case Enums::Ring:     Area = PI * (OuterRadius * OuterRadius) - PI * (InnerRadius * InnerRadius); break;
        case Enums::Box:      Area = (2.0f * Size[0] * Size[1]) + (2.0f * Size[0] * Size[2]) + (2.0f * Size[1] * Size[2]); break;
        case Enums::Sphere:   Area = 4.0f * PI * (OuterRadius * OuterRadius); break;
        case Enums::Cylinder: Area = (2.0f * PI * (OuterRadius * OuterRadius)) + (2.0f * PI * OuterRadius * Size[2]); break;
        default: break;
    }
}

Procedural::Procedural() : Type(Enums::Uniform) {}

ErTracer::ErTracer() : ErBindable(), NoIterations(0), VolumeID(0) {}

void BindTracer(const ErTracer& T, const bool& Bind)
{
    er_tracer_desc d;
    er_tracer_desc_init(&d);
    d.id = T.ID;
    d.enabled = T.Enabled ? 1 : 0;
    to_plf(T.Opacity1D.PLF, d.opacity);
    to_ctf(T.Diffuse1D, d.diffuse);
    to_ctf(T.Specular1D, d.specular);
    to_plf(T.Glossiness1D.PLF, d.glossiness);
    to_ctf(T.Emission1D, d.emission);
    const Camera& c = T.Camera;
    d.camera.film_size[0] = c.FilmSize[0];
    d.camera.film_size[1] = c.FilmSize[1];
    for (int i = 0; i < 3; ++i) {
        d.camera.pos[i] = c.Pos[i];
        d.camera.target[i] = c.Target[i];
        d.camera.up[i] = c.Up[i];
    }
    d.camera.focal_distance = c.FocalDistance;
    d.camera.aperture_size = c.ApertureSize;
    d.camera.clip_near = c.ClipNear;
    d.camera.clip_far = c.ClipFar;
    d.camera.exposure = c.Exposure;
    d.camera.gamma = c.Gamma;
    d.camera.fov = c.FOV;
    const RenderSettings& r = T.RenderSettings;
    er_render_settings& s = d.settings;
    s.step_factor_primary = r.Traversal.StepFactorPrimary;
    s.step_factor_shadow = r.Traversal.StepFactorShadow;
    s.shadows = r.Traversal.Shadows ? 1 : 0;
    s.max_shadow_distance = r.Traversal.MaxShadowDistance;
    s.shading_type = r.Shading.Type;
    s.density_scale = r.Shading.DensityScale;
    s.opacity_modulated = r.Shading.OpacityModulated ? 1 : 0;
    s.gradient_computation = r.Shading.GradientComputation;
    s.gradient_threshold = r.Shading.GradientThreshold;
    s.gradient_factor = r.Shading.GradientFactor;
    s.estimator = (int)r.Estimator;
    s.sampler = (int)r.Sampler;
    s.bilateral = r.Bilateral ? 1 : 0;
    s.bilateral_radius = r.BilateralRadius;
    s.bilateral_sigma_spatial = r.BilateralSigmaSpatial;
    s.bilateral_sigma_range = r.BilateralSigmaRange;
    d.no_iterations = T.NoIterations;
    d.volume_id = T.VolumeID;
    to_indices(T.LightIDs, d.light_ids, d.light_count);
    to_indices(T.ObjectIDs, d.object_ids, d.object_count);
    to_indices(T.ClippingObjectIDs, d.clipping_ids, d.clipping_count);
    const int st = er_bind_tracer(&d, Bind ? 1 : 0);
    T.ID = d.id;
    check(st);
}

void BindVolume(const ErVolume& V, const bool& Bind)
{
    er_volume_desc d;
    er_volume_desc_init(&d);
    d.id = V.ID;
    d.enabled = V.Enabled ? 1 : 0;
    for (int i = 0; i < 3; ++i) {
        d.resolution[i] = V.Voxels.Resolution[i];
        d.spacing[i] = V.Spacing[i];
    }
    d.normalize_size = V.NormalizeSize ? 1 : 0;
    d.voxels_dirty = V.Voxels.Dirty ? 1 : 0;
    d.voxels = V.Voxels.GetData();
    const int st = er_bind_volume(&d, Bind ? 1 : 0);
    V.ID = d.id;
    if (st == ER_OK && Bind) V.Voxels.Dirty = false;
    check(st);
}

void BindLight(const ErLight& L, const bool& Bind)
{
    er_light_desc d;
    er_light_desc_init(&d);
    d.id = L.ID;
    d.enabled = L.Enabled ? 1 : 0;
    d.visible = L.Visible ? 1 : 0;
    d.texture_id = L.TextureID;
    d.multiplier = L.Multiplier;
    d.unit = (int)L.Unit;
    to_shape(L.Shape, d.shape);
    const int st = er_bind_light(&d, Bind ? 1 : 0);
    L.ID = d.id;
    check(st);
}

void BindObject(const ErObject& O, const bool& Bind)
{
    er_object_desc d;
    er_object_desc_init(&d);
    d.id = O.ID;
    d.enabled = O.Enabled ? 1 : 0;
    d.diffuse_texture_id = O.DiffuseTextureID;
    d.specular_texture_id = O.SpecularTextureID;
    d.glossiness_texture_id = O.GlossinessTextureID;
    d.ior = O.Ior;
    to_shape(O.Shape, d.shape);
    const int st = er_bind_object(&d, Bind ? 1 : 0);
    O.ID = d.id;
    check(st);
}

void BindClippingObject(const ErClippingObject& C, const bool& Bind)
{
    er_clipping_desc d;
    er_clipping_desc_init(&d);
    d.id = C.ID;
    d.enabled = C.Enabled ? 1 : 0;
    d.invert = C.Invert ? 1 : 0;
    to_shape(C.Shape, d.shape);
    const int st = er_bind_clipping_object(&d, Bind ? 1 : 0);
    C.ID = d.id;
    check(st);
}

void BindTexture(const ErTexture& T, const bool& Bind)
{
    er_texture_desc d;
    er_texture_desc_init(&d);
    d.id = T.ID;
    d.enabled = T.Enabled ? 1 : 0;
    d.type = (int)T.Type;
    d.output_level = T.OutputLevel;
    d.bitmap_id = T.BitmapID;
    for (int i = 0; i < 2; ++i) {
        d.offset[i] = T.Offset[i];
        d.repeat[i] = T.Repeat[i];
        d.flip[i] = T.Flip[i];
    }
    d.procedural.type = (int)T.Procedural.Type;
    for (int i = 0; i < 3; ++i) {
        d.procedural.uniform_color[i] = T.Procedural.UniformColor[i];
        d.procedural.checker_color1[i] = T.Procedural.CheckerColor1[i];
        d.procedural.checker_color2[i] = T.Procedural.CheckerColor2[i];
    }
    to_ctf(T.Procedural.Gradient, d.procedural.gradient);
    const int st = er_bind_texture(&d, Bind ? 1 : 0);
    T.ID = d.id;
    check(st);
}

void BindBitmap(const ErBitmap& B, const bool& Bind)
{
    er_bitmap_desc d;
    er_bitmap_desc_init(&d);
    d.id = B.ID;
    d.enabled = B.Enabled ? 1 : 0;
    d.resolution[0] = B.Pixels.Resolution[0];
    d.resolution[1] = B.Pixels.Resolution[1];
    d.pixels_dirty = B.Pixels.Dirty ? 1 : 0;
    d.pixels = reinterpret_cast<const uint8_t*>(B.Pixels.GetData());
    const int st = er_bind_bitmap(&d, Bind ? 1 : 0);
    B.ID = d.id;
    if (st == ER_OK && Bind) B.Pixels.Dirty = false;
    check(st);
}

void RenderEstimate(int TracerID) { check(er_render_estimate(TracerID)); }

/* The reference writes W*H*4 bytes without a size argument; the caller guarantees the room. */
void GetEstimate(int TracerID, unsigned char* pData) { check(er_get_estimate(TracerID, pData, SIZE_MAX)); }

void GetAutoFocusDistance(int TracerID, int FilmU, int FilmV, float& AutoFocusDistance)
{
    check(er_get_autofocus_distance(TracerID, FilmU, FilmV, &AutoFocusDistance));
}

void GetNoIterations(int TracerID, int& NoIterations) { check(er_get_no_iterations(TracerID, &NoIterations)); }

}  // namespace ExposureRender
