// Scene scripts (.ilm) for the GPU path: the reference's lexer and parse
// tree (Core/System/Lexer.cpp:27-114, Core/Scene/EnvironmentLoader.cpp:27-213)
// and the resolution of the parsed sections into one renderable job
// (EnvironmentLoader.cpp:215-931): shapes, materials, lights, cameras,
// filters, samplers, devices, integrators, renderers and the Environment /
// Scene / Space / Primitives block.
//
// Resolution is host-only (no device is touched), so a script can be loaded
// and inspected without a GPU; RenderJob::Render draws it with the GPU path.
// Shapes: Quad, Triangle, WavefrontModel and PersistentMesh (geometry and
// group ids; the flattened space gets one tree). What the GPU path cannot
// draw (spheres, mirror and glass materials, point,
// spot, directional and environment lights, the thin-lens camera, integrators
// other than PathTracing) is recorded when its section is parsed and becomes
// an error only if the environment uses it, as the reference reports a failed
// instance and carries on.
#pragma once

#include <istream>
#include <map>
#include <string>
#include <vector>

#include "ilm_plugin.hpp"

namespace IlluminaGpu {

struct LexerToken {
    enum TokenType { Undefined, LeftCurly, RightCurly, PropertyName, PropertyValue, EndOfStream };
    TokenType Type = Undefined;
    std::string Value;
};

// Tokens: '{', '}', a name ([A-Za-z_][A-Za-z0-9_]*, ended by a blank; other
// characters inside a name are dropped), "= value" up to the end of the line
// (trimmed), '#' comments to the end of the line. A stream that ends inside
// a token throws, as the reference does.
class Lexer {
public:
    explicit Lexer(std::istream* in) : m_in(in) {}
    bool ReadToken(LexerToken& token);
private:
    std::istream* m_in;
};

struct ParseNode {
    enum NodeType { Internal, Leaf };
    NodeType Type = Internal;
    std::string Name, Value;
    std::vector<ParseNode*> Children;
    bool FindByName(const std::string& name, std::vector<ParseNode*>& out) const;
    // Leaf children as name -> value (the last of equal names wins).
    bool GetArgumentMap(ArgumentMap& args) const;
};

// A scene ready for the GPU path: the space's primitives flattened to world
// space (BasicSpace's per-primitive transforms applied, ilm_space_*),
// per-triangle materials and lights, the camera and the renderer's settings.
struct RenderJob {
    ilm_camera Camera{};
    int Width = 640, Height = 480, Samples = 1;
    int RayDepth = 6;
    float Epsilon = 1e-4f;
    std::string Filename = "result.ppm";
    int Format = ILM_IMAGE_PPM;
    int Builder = ILM_BUILDER_SAH;
    std::vector<float> Positions, Normals, Materials, Lights;
    std::vector<int32_t> Indices, TriMaterial, TriLight;
    std::vector<std::string> PrimitiveIds;
    std::vector<int64_t> PrimitiveFirstTriangle;  // plus one entry past the end
    std::vector<std::string> LightIds;            // in the Scene's Lights order

    // Renders Samples spp on the GPU and returns the per-pixel mean radiance
    // in the order the image device writes it (ImageDevice mirrors the
    // frame, Device/ImageDevice.cpp:137-148); writes Filename when asked.
    void Render(uint32_t seed, std::vector<float>& image, bool write) const;
};

class EnvironmentLoader {
public:
    // EnvironmentLoader::Load: false when the file cannot be opened. Throws
    // Exception* for a script whose environment the GPU path cannot render.
    bool Load(const std::string& filename);
    // Parses and resolves a script from a stream; relative model file names
    // are tried as given, then against base_dir.
    void Load(std::istream& in, const std::string& base_dir);
    const ParseNode& Root() const { return m_root; }
    const RenderJob& Job() const { return m_job; }
    const std::vector<std::string>& Warnings() const { return m_warnings; }
    ~EnvironmentLoader();

private:
    struct Shape {
        std::vector<float> pos, nrm;
        std::vector<int32_t> idx, group;  // group: per-triangle material group id (OBJ usemtl order)
        int builder = ILM_BUILDER_SAH;
        std::string unsupported;
    };
    struct Material {
        std::vector<float> records;  // 8 floats per member; one member unless a group
        std::vector<std::string> member_unsupported;  // per member, "" when drawable
        bool group = false;
        std::string unsupported;
    };
    struct Light {
        float power[3] = {0, 0, 0};
        std::string shape, unsupported;
    };
    struct Camera {
        ilm_camera cam{};
        std::string unsupported;
    };
    struct Integrator {
        int depth = 6;
        float eps = 1e-4f;
        std::string unsupported;
    };
    struct Renderer {
        int samples = 1;
        std::string integrator, device, filter, unsupported;
    };
    struct Device {
        int width = 640, height = 480;
        std::string format = "PPM", filename = "result.ppm", unsupported;
    };

    void Parse(std::istream& in);
    bool NodeList(const std::string& category, const std::string& instance, std::vector<ParseNode*>& out) const;
    template <class T>
    void Each(const std::string& category, const std::string& instance, const char* label,
              std::map<std::string, T>& into, T (EnvironmentLoader::*make)(const std::string&, const ArgumentMap&));
    Shape MakeShape(const std::string& type, const ArgumentMap& args);
    Material MakeMaterial(const std::string& type, const ArgumentMap& args);
    Light MakeLight(const std::string& type, const ArgumentMap& args);
    Camera MakeCamera(const std::string& type, const ArgumentMap& args);
    Integrator MakeIntegrator(const std::string& type, const ArgumentMap& args);
    Renderer MakeRenderer(const std::string& type, const ArgumentMap& args);
    Device MakeDevice(const std::string& type, const ArgumentMap& args);
    std::string MakeName(const std::string& type, const ArgumentMap& args);
    void ParseEnvironment();
    void Warn(const std::string& w) { m_warnings.push_back(w); }
    std::string Resolve(const std::string& file) const;

    ParseNode m_root;
    std::vector<ParseNode*> m_nodes;  // owned
    std::string m_base;
    std::map<std::string, Shape> m_shapes;
    std::map<std::string, Material> m_materials;
    std::map<std::string, Light> m_lights;
    std::map<std::string, Camera> m_cameras;
    std::map<std::string, Integrator> m_integrators;
    std::map<std::string, Renderer> m_renderers;
    std::map<std::string, Device> m_devices;
    std::map<std::string, std::string> m_filters, m_samplers, m_textures;  // id -> unsupported reason ("" = usable)
    RenderJob m_job;
    std::vector<std::string> m_warnings;
};

}  // namespace IlluminaGpu
