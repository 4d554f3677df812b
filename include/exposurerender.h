/*
 * exposurerender.h - the C++ API of the reference (Source/exposurerender.h:29-42) over the
 * C-ABI in er_capi.h.
 *
 * Same namespace, function names, signatures, descriptor class names and public field names
 * as the reference, so an application written against the reference compiles against this
 * header unchanged. Semantics kept:
 *   - Bind* writes the assigned ID into the caller's const descriptor (mutable ID,
 *     Source/erbindable.h:65; Source/list.cuh:69-76); a rebind with the same ID updates in
 *     place; Bind == false unbinds.
 *   - Failures throw ExposureRender::Exception{Level, Message} (Source/exception.h:27-55):
 *     Warning for an unknown ID, Error for a device failure or a bad argument.
 *   - Every call is synchronous (Source/wrapper.cuh:53-135): RenderEstimate returns with the
 *     display image final.
 *   - BindVoxels / BindPixels deep-copy the caller's array (Source/buffer3d.h:166-198); the
 *     upload is gated by the buffer's Dirty flag, which Bind clears (buffer3d.h:52-65).
 * Additions: RenderSettings carries the estimator mode, free-flight sampler and bilateral
 * display filter of the C-ABI (defaults reproduce the reference), and GetNoIterations
 * returns the real pass count (a no-op in the reference, core.cu:150-153).
 */
#ifndef EXPOSURE_RENDER_H
#define EXPOSURE_RENDER_H

#include <cfloat>
#include <cstring>
#include <vector>

#include "er_capi.h"

#if defined(__GNUC__) || defined(__clang__)
#define EXPOSURE_RENDER_DLL __attribute__((visibility("default")))
#else
#define EXPOSURE_RENDER_DLL
#endif

#define MAX_NO_TF_NODES ER_MAX_NO_TF_NODES
#define MAX_CHAR_SIZE 256

namespace ExposureRender
{

/* Source/enums.h:24-113 (same enumerators, same values) */
namespace Enums
{
enum MemoryType { Host, Device };
enum ProceduralType { Uniform = 0, Checker, Gradient };
enum TextureType { Procedural = 0, Bitmap };
enum ShapeType { Plane = 0, Disk, Ring, Box, Sphere, Cylinder, Cone };
enum ShadingMode { BrdfOnly = 0, PhaseFunctionOnly, Hybrid, Modulation, Threshold, GradientMagnitude };
enum GradientMode { ForwardDifferences = 0, CentralDifferences, Filtered };
enum ExceptionLevel { Info = 0, Warning, Error, Fatal };
enum EmissionUnit { Power = 0, Lux, Intensity };
/* additions of this implementation (er_capi.h ER_ESTIMATOR_* / ER_SAMPLER_*) */
enum EstimatorMode { Compat = ER_ESTIMATOR_COMPAT, Full = ER_ESTIMATOR_FULL };
enum FreeFlightSampler { RayMarch = ER_SAMPLER_RAYMARCH, Woodcock = ER_SAMPLER_WOODCOCK };
}

/* Source/exception.h:27-55 */
class EXPOSURE_RENDER_DLL Exception
{
public:
    Exception(const Enums::ExceptionLevel& Level, const char* pMessage = "") : Level(Level)
    {
        std::strncpy(Message, pMessage ? pMessage : "", MAX_CHAR_SIZE - 1);
        Message[MAX_CHAR_SIZE - 1] = '\0';
    }
    Enums::ExceptionLevel Level;
    char Message[MAX_CHAR_SIZE];
};

/* Fixed-size vectors with the element access the descriptors use (Source/vector.h). */
template<class T, int N>
class Vec
{
public:
    Vec() { for (int i = 0; i < N; ++i) D[i] = T(0); }
    explicit Vec(T v) { for (int i = 0; i < N; ++i) D[i] = v; }
    Vec(T x, T y) { static_assert(N == 2, "2 components"); D[0] = x; D[1] = y; }
    Vec(T x, T y, T z) { static_assert(N == 3, "3 components"); D[0] = x; D[1] = y; D[2] = z; }
    T& operator[](int i) { return D[i]; }
    const T& operator[](int i) const { return D[i]; }
    T D[N];
};
typedef Vec<int, 2> Vec2i;
typedef Vec<float, 2> Vec2f;
typedef Vec<int, 3> Vec3i;
typedef Vec<float, 3> Vec3f;

/* CIE XYZ colour (Source/color.h) and 8-bit RGBA pixel */
class ColorXYZf : public Vec<float, 3>
{
public:
    ColorXYZf() : Vec<float, 3>() {}
    explicit ColorXYZf(float v) : Vec<float, 3>(v) {}
    ColorXYZf(float x, float y, float z) : Vec<float, 3>(x, y, z) {}
};
struct ColorRGBAuc { unsigned char D[4]; };

/* Row-major 4x4 matrix NN[row][col] (Source/matrix.h); identity by default */
class Matrix44
{
public:
    Matrix44() { for (int r = 0; r < 4; ++r) for (int c = 0; c < 4; ++c) NN[r][c] = r == c ? 1.0f : 0.0f; }
    float NN[4][4];
};

/* 256 ints + Count (Source/vector.h:554-579) */
class Indices
{
public:
    Indices() : Count(0) { std::memset(IDs, 0, sizeof(IDs)); }
    int& operator[](int i) { return IDs[i]; }
    const int& operator[](int i) const { return IDs[i]; }
    int IDs[ER_MAX_INDICES];
    int Count;
};

/* PiecewiseLinearFunction (Source/plf.h:26-92, Source/pf.h:44-71) */
template<int Size = MAX_NO_TF_NODES>
class PiecewiseLinearFunction
{
public:
    PiecewiseLinearFunction() : NodeRange(FLT_MAX, FLT_MIN), Count(0) {}
    void AddNode(const float& P, const float& V)
    {
        if (Count + 1 >= MAX_NO_TF_NODES) return;
        Position[Count] = P;
        Value[Count] = V;
        if (P < NodeRange[0]) NodeRange[0] = P;
        if (P > NodeRange[1]) NodeRange[1] = P;
        Count++;
    }
    Vec2f NodeRange;
    Vec<float, Size> Position;
    Vec<float, Size> Value;
    int Count;
};

/* Source/transferfunction.h:27-154 */
class ScalarNode
{
public:
    ScalarNode() : Position(0.0f), Value(0.0f) {}
    ScalarNode(float Position, float Value) : Position(Position), Value(Value) {}
    float Position;
    float Value;
};

class ColorNode
{
public:
    ColorNode() {}
    ColorNode(float Position, float X, float Y, float Z)
    {
        ScalarNodes[0] = ScalarNode(Position, X);
        ScalarNodes[1] = ScalarNode(Position, Y);
        ScalarNodes[2] = ScalarNode(Position, Z);
    }
    ScalarNode ScalarNodes[3];
};

class ScalarTransferFunction1D
{
public:
    void AddNode(const ScalarNode& Node) { PLF.AddNode(Node.Position, Node.Value); }
    PiecewiseLinearFunction<MAX_NO_TF_NODES> PLF;
};

class ColorTransferFunction1D
{
public:
    void AddNode(const ColorNode& Node)
    {
        for (int i = 0; i < 3; ++i) PLF[i].AddNode(Node.ScalarNodes[i].Position, Node.ScalarNodes[i].Value);
    }
    PiecewiseLinearFunction<MAX_NO_TF_NODES> PLF[3];
};

/* Camera (Source/camera.h:26-109): the caller-set fields; N/U/V/Screen are derived at bind */
class EXPOSURE_RENDER_DLL Camera
{
public:
    Camera();
    Vec2i FilmSize;
    Vec3f Pos;
    Vec3f Target;
    Vec3f Up;
    float FocalDistance;
    float ApertureSize;
    float ClipNear;
    float ClipFar;
    float Exposure;
    float Gamma;
    float FOV;
};

/* RenderSettings (Source/rendersettings.h:27-131) */
class EXPOSURE_RENDER_DLL RenderSettings
{
public:
    class TraversalSettings
    {
    public:
        float StepFactorPrimary;
        float StepFactorShadow;
        bool Shadows;
        float MaxShadowDistance;
    };
    class ShadingSettings
    {
    public:
        int Type;
        float DensityScale;
        bool OpacityModulated;
        int GradientComputation;
        float GradientThreshold;
        float GradientFactor;
    };
    RenderSettings();
    TraversalSettings Traversal;
    ShadingSettings Shading;
    /* additions */
    Enums::EstimatorMode Estimator;
    Enums::FreeFlightSampler Sampler;
    bool Bilateral;
    int BilateralRadius;
    float BilateralSigmaSpatial;
    float BilateralSigmaRange;
};

/* Shape (Source/shape.h:57-112): the caller supplies both TM and InvTM; Area is derived */
class EXPOSURE_RENDER_DLL Shape
{
public:
    Shape();
    void Update();
    Matrix44 TM;
    Matrix44 InvTM;
    bool OneSided;
    Enums::ShapeType Type;
    Vec3f Size;
    float Area;
    float InnerRadius;
    float OuterRadius;
};

/* Procedural (Source/procedural.h:27-60) */
class EXPOSURE_RENDER_DLL Procedural
{
public:
    Procedural();
    Enums::ProceduralType Type;
    ColorXYZf UniformColor;
    ColorXYZf CheckerColor1;
    ColorXYZf CheckerColor2;
    ColorTransferFunction1D Gradient;
};

/* Host buffers with the dirty-gated upload of Source/buffer3d.h:52-65 / buffer2d.h */
template<class T>
class Buffer3D
{
public:
    Buffer3D(const Enums::MemoryType& = Enums::Host, const char* = "") : Dirty(false) {}
    void Set(const Enums::MemoryType&, const Vec3i& R, const T* pData)
    {
        Resolution = R;
        const size_t n = (size_t)(R[0] > 0 ? R[0] : 0) * (size_t)(R[1] > 0 ? R[1] : 0) * (size_t)(R[2] > 0 ? R[2] : 0);
        Data.assign(pData, pData + (pData ? n : 0));
        Dirty = true;
    }
    const Vec3i& GetResolution() const { return Resolution; }
    const T* GetData() const { return Data.empty() ? nullptr : Data.data(); }
    Vec3i Resolution;
    std::vector<T> Data;
    mutable bool Dirty;
};

template<class T>
class Buffer2D
{
public:
    Buffer2D(const Enums::MemoryType& = Enums::Host, const char* = "") : Dirty(false) {}
    void Set(const Enums::MemoryType&, const Vec2i& R, const T* pData)
    {
        Resolution = R;
        const size_t n = (size_t)(R[0] > 0 ? R[0] : 0) * (size_t)(R[1] > 0 ? R[1] : 0);
        Data.assign(pData, pData + (pData ? n : 0));
        Dirty = true;
    }
    const Vec2i& GetResolution() const { return Resolution; }
    const T* GetData() const { return Data.empty() ? nullptr : Data.data(); }
    Vec2i Resolution;
    std::vector<T> Data;
    mutable bool Dirty;
};

/* ErBindable (Source/erbindable.h:27-67) */
class EXPOSURE_RENDER_DLL ErBindable
{
public:
    ErBindable() : ID(-1), Enabled(true), Dirty(false) {}
    virtual ~ErBindable() {}
    mutable int ID;
    bool Enabled;
    bool Dirty;
};

/* ErTracer (Source/ertracer.h:30-117) */
class EXPOSURE_RENDER_DLL ErTracer : public ErBindable
{
public:
    ErTracer();
    ScalarTransferFunction1D Opacity1D;
    ColorTransferFunction1D Diffuse1D;
    ColorTransferFunction1D Specular1D;
    ScalarTransferFunction1D Glossiness1D;
    ColorTransferFunction1D Emission1D;
    ExposureRender::Camera Camera;
    ExposureRender::RenderSettings RenderSettings;
    int NoIterations;
    int VolumeID;
    Indices LightIDs;
    Indices ObjectIDs;
    Indices ClippingObjectIDs;
};

/* ErVolume (Source/ervolume.h:27-74) */
class EXPOSURE_RENDER_DLL ErVolume : public ErBindable
{
public:
    ErVolume() : ErBindable(), Voxels(Enums::Host, "Host Voxels"), NormalizeSize(false), Spacing(1.0f) {}
    void BindVoxels(const Vec3i& Resolution, const Vec3f& Spacing, unsigned short* Voxels, const bool& NormalizeSize = false)
    {
        this->Voxels.Set(Enums::Host, Resolution, Voxels);
        this->NormalizeSize = NormalizeSize;
        this->Spacing = Spacing;
    }
    Buffer3D<unsigned short> Voxels;
    bool NormalizeSize;
    Vec3f Spacing;
};

/* ErLight (Source/erlight.h:27-66) */
class EXPOSURE_RENDER_DLL ErLight : public ErBindable
{
public:
    ErLight() : ErBindable(), Visible(false), TextureID(-1), Multiplier(0.0f), Unit(Enums::Power) {}
    bool Visible;
    ExposureRender::Shape Shape;
    int TextureID;
    float Multiplier;
    Enums::EmissionUnit Unit;
};

/* ErObject (Source/erobject.h:27-67) */
class EXPOSURE_RENDER_DLL ErObject : public ErBindable
{
public:
    ErObject() : ErBindable(), DiffuseTextureID(-1), SpecularTextureID(-1), GlossinessTextureID(-1), Ior(5.0f) {}
    ExposureRender::Shape Shape;
    int DiffuseTextureID;
    int SpecularTextureID;
    int GlossinessTextureID;
    float Ior;
};

/* ErClippingObject (Source/erclippingobject.h:27-58) */
class EXPOSURE_RENDER_DLL ErClippingObject : public ErBindable
{
public:
    ErClippingObject() : ErBindable(), Invert(false) {}
    ExposureRender::Shape Shape;
    bool Invert;
};

/* ErTexture (Source/ertexture.h:27-82) */
class EXPOSURE_RENDER_DLL ErTexture : public ErBindable
{
public:
    ErTexture() : ErBindable(), Type(Enums::Procedural), OutputLevel(1.0f), BitmapID(-1), Offset(0.0f), Repeat(0.0f) {}
    Enums::TextureType Type;
    float OutputLevel;
    int BitmapID;
    ExposureRender::Procedural Procedural;
    Vec2f Offset;
    Vec2f Repeat;
    Vec2i Flip;
};

/* ErBitmap (Source/erbitmap.h:28-61) */
class EXPOSURE_RENDER_DLL ErBitmap : public ErBindable
{
public:
    ErBitmap() : ErBindable(), Pixels(Enums::Host, "Host Pixels") {}
    void BindPixels(const Vec2i& Resolution, ColorRGBAuc* Pixels) { this->Pixels.Set(Enums::Host, Resolution, Pixels); }
    Buffer2D<ColorRGBAuc> Pixels;
};

/* Source/exposurerender.h:32-42 */
EXPOSURE_RENDER_DLL void BindTracer(const ErTracer& Tracer, const bool& Bind = true);
EXPOSURE_RENDER_DLL void BindVolume(const ErVolume& Volume, const bool& Bind = true);
EXPOSURE_RENDER_DLL void BindLight(const ErLight& Light, const bool& Bind = true);
EXPOSURE_RENDER_DLL void BindObject(const ErObject& Object, const bool& Bind = true);
EXPOSURE_RENDER_DLL void BindClippingObject(const ErClippingObject& ClippingObject, const bool& Bind = true);
EXPOSURE_RENDER_DLL void BindTexture(const ErTexture& Texture, const bool& Bind = true);
EXPOSURE_RENDER_DLL void BindBitmap(const ErBitmap& Bitmap, const bool& Bind = true);
EXPOSURE_RENDER_DLL void RenderEstimate(int TracerID);
EXPOSURE_RENDER_DLL void GetEstimate(int TracerID, unsigned char* pData);
EXPOSURE_RENDER_DLL void GetAutoFocusDistance(int TracerID, int FilmU, int FilmV, float& AutoFocusDistance);
EXPOSURE_RENDER_DLL void GetNoIterations(int TracerID, int& NoIterations);

}

#endif /* EXPOSURE_RENDER_H */
