// C++ host layer above the C ABI, mirroring the reference renderer's plug-in
// interface for this path: Factory / FactoryManager / ArgumentMap /
// IPlugIn (Core/System/Factory.h:15-21, FactoryManager.inl:63-177,
// ArgumentMap.h, PlugIn.h:18-38), a mesh shape "GpuBVHMesh" in the place of
// BVHMesh (Core/Shape/BVHMesh.h:29-106) and an integrator "GpuPathTracing"
// in the place of PathIntegrator (Core/Integrator/PathIntegrator.h:28-29).
// Errors are thrown as heap-allocated exceptions, as the reference does
// (throw new Exception(...), Core/System/FactoryManager.inl:69).
#pragma once

#include <stdint.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/ilm_gpu.h"

namespace IlluminaGpu {

class Exception {
public:
    explicit Exception(const std::string& m) : m_message(m) {}
    const std::string& GetMessage() const { return m_message; }
private:
    std::string m_message;
};

inline void Check(int rc) {
    if (rc != ILM_OK) throw new Exception(ilm_last_error());
}

struct Vector3 {
    float X = 0, Y = 0, Z = 0;
    Vector3() {}
    Vector3(float x, float y, float z) : X(x), Y(y), Z(z) {}
};

// Name -> value text, read as the reference reads it (Core/System/ArgumentMap.h):
// numbers and strings through a stream (a string is its first
// whitespace-delimited word), vectors as "{x, y, z}" (one separator
// character between numbers), a Matrix3x3 as nine numbers in storage order,
// and lists as "{a, b, ...}" split at '{', ',' and '}' (list items are
// trimmed here; the reference keeps their leading blanks).
class ArgumentMap {
public:
    ArgumentMap() {}
    // "Key = Value; Key = Value" (ArgumentMap::Initialise(string), :53-73)
    explicit ArgumentMap(const std::string& text);
    explicit ArgumentMap(const std::map<std::string, std::string>& args) : m_args(args) {}
    bool ContainsArgument(const std::string& name) const { return m_args.count(name) != 0; }
    bool GetArgument(const std::string& name, int& v) const;
    bool GetArgument(const std::string& name, float& v) const;
    bool GetArgument(const std::string& name, std::string& v) const;
    bool GetArgument(const std::string& name, Vector3& v) const;
    bool GetArgument(const std::string& name, float (&m)[9]) const;
    bool GetArgument(const std::string& name, std::vector<std::string>& v) const;
    bool GetArgument(const std::string& name, std::vector<Vector3>& v) const;
private:
    std::map<std::string, std::string> m_args;
};

// Geometry/Ray.h:17-54: Min is kept but, as in the reference triangle test,
// hits are accepted for 1e-5 <= t <= Max - 1e-5.
struct Ray {
    Vector3 Origin, Direction;
    float Min = 0.0f, Max = 3.402823466e38f;
};

// The subset of DifferentialSurface (Core/Geometry/Intersection.h) the path fills.
struct DifferentialSurface {
    float Distance = 0;
    Vector3 Point;
    int TriangleIndex = -1;  // position in the mesh's triangle list (the reference's hit index)
    float Beta = 0, Gamma = 0;
};

class IShape {
public:
    virtual ~IShape() {}
    virtual bool Compile() = 0;
    virtual bool Intersects(const Ray& ray, DifferentialSurface& surface) = 0;
    virtual bool Intersects(const Ray& ray) = 0;
};

class GpuBVHMesh : public IShape {
public:
    GpuBVHMesh(ilm_ctx* ctx, int maxLeafObjects, int maxTreeDepth, int builder);
    ~GpuBVHMesh();
    // ITriangleMesh::AddVertex / AddIndexedTriangle (Core/Shape/TriangleMesh.h:41-94)
    void AddVertex(const Vector3& p);
    void AddIndexedTriangle(int v0, int v1, int v2);
    bool Compile();  // BVHMesh::Compile (Core/Shape/BVHMesh.cpp:78-107)
    // Per-ray IShape::Intersects: a one-ray batch (correct, latency-bound).
    bool Intersects(const Ray& ray, DifferentialSurface& surface);
    bool Intersects(const Ray& ray);
    // Batched closest / any hit on host arrays (uploads, traces, downloads).
    void IntersectBatch(const std::vector<Ray>& rays, std::vector<DifferentialSurface>& out);
    void OccludedBatch(const std::vector<Ray>& rays, std::vector<uint8_t>& out);
    ilm_mesh* Handle() const { return m_mesh; }
    const std::vector<float>& Positions() const { return m_pos; }
    const std::vector<int32_t>& Indices() const { return m_idx; }
private:
    ilm_ctx* m_ctx;
    ilm_mesh* m_mesh = nullptr;
    int m_maxLeaf, m_maxDepth, m_builder;
    std::vector<float> m_pos;
    std::vector<int32_t> m_idx;
    void* m_stage = nullptr;  // one-ray staging (ray + hit)
};

// The subset of Intersection (Core/Geometry/Intersection.h) a space fills:
// the surface, and which primitive of the space's list was hit.
struct Intersection {
    DifferentialSurface Surface;  // TriangleIndex is the triangle within the primitive
    int PrimitiveIndex = -1;
};

// A primitive of a space: GeometricPrimitive (Scene/GeometricPrimitive.h) =
// shape + WorldTransform (rotation as the reference's Matrix3x3 elements in
// storage order, scaling, translation) + material; EmissivePrimitive adds a
// light (Light >= 0).
struct Primitive {
    GpuBVHMesh* Shape = nullptr;
    float Rotation[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    float Scaling[3] = {1, 1, 1};
    float Translation[3] = {0, 0, 0};
    int Material = 0, Light = -1;
};

// ISpace (Space/Space.h:31-43) without the exclude-primitive overloads.
class ISpace {
public:
    virtual ~ISpace() {}
    std::vector<Primitive> PrimitiveList;
    virtual bool Initialise() = 0;
    virtual void Shutdown() = 0;
    virtual bool Update() = 0;
    virtual bool Build() = 0;
    virtual bool Intersects(const Ray& ray, Intersection& intersection) = 0;
    virtual bool Intersects(const Ray& ray) = 0;
};

// BasicSpace (Space/BasicSpace.cpp:58-114) on the GPU: Build flattens the
// primitive list into one world-space BVH (ilm_space_*).
class GpuSpace : public ISpace {
public:
    explicit GpuSpace(ilm_ctx* ctx);
    ~GpuSpace();
    bool Initialise() override { return true; }
    void Shutdown() override {}
    bool Update() override { return true; }
    bool Build() override;
    bool Intersects(const Ray& ray, Intersection& intersection) override;
    bool Intersects(const Ray& ray) override;
    void IntersectBatch(const std::vector<Ray>& rays, std::vector<Intersection>& out);
    ilm_space* Handle() const { return m_space; }
private:
    ilm_ctx* m_ctx;
    ilm_space* m_space = nullptr;
    ilm_mesh* m_mesh = nullptr;
    void* m_stage = nullptr;  // one ray, hit, primitive and triangle index
};

// Camera arguments as CameraFactories.h:45-63 reads them.
struct CameraArgs {
    Vector3 Position, Target, Up = Vector3(0, 1, 0);
    float Fov = 60.0f, Aspect = 1.0f;
};

class IIntegrator {
public:
    virtual ~IIntegrator() {}
};

// PathIntegrator arguments RayDepth (default 6), ShadowRays (1), Epsilon
// (1e-4) (Core/Integrator/IntegratorFactories.h:43-52). Shading data comes
// with the scene; RenderRegion renders whole images on the GPU and returns
// per-pixel mean radiance (row-major, 3 floats per pixel).
class GpuPathTracing : public IIntegrator {
public:
    GpuPathTracing(ilm_ctx* ctx, int rayDepth, float epsilon);
    ~GpuPathTracing();
    void Initialise(GpuBVHMesh* mesh, const std::vector<float>& vertexNormals, const std::vector<int32_t>& triMaterial,
                    const std::vector<int32_t>& triLight, const std::vector<float>& materials,
                    const std::vector<float>& lights);
    void RenderRegion(const CameraArgs& cam, int width, int height, int spp, uint32_t seed, std::vector<float>& image);
private:
    ilm_ctx* m_ctx;
    ilm_scene* m_scene = nullptr;
    int m_rayDepth;
    float m_epsilon;
};

// Factory<T> / FactoryManager<T> (Core/System/Factory.h:15-21, FactoryManager.inl:63-177)
template <class T>
class Factory {
public:
    virtual ~Factory() {}
    virtual T* CreateInstance(ArgumentMap& args) = 0;
};

// Same error cases and messages as the reference manager: duplicate factory
// or instance names, unknown factories and failed creation throw.
template <class T>
class FactoryManager {
public:
    ~FactoryManager() {
        for (auto& kv : m_instances) delete kv.second;
    }
    void RegisterFactory(const std::string& name, Factory<T>* factory) {
        if (m_factories.count(name))
            throw new Exception("Cannot register factory. Another factory is already registered with the same name!");
        m_factories[name] = factory;
    }
    void UnregisterFactory(const std::string& name) {
        if (!m_factories.count(name)) throw new Exception("Cannot unregister factory. No factory found!");
        m_factories.erase(name);
    }
    bool QueryFactory(const std::string& name) const { return m_factories.count(name) != 0; }
    T* CreateInstance(const std::string& factory, const std::string& instanceName, ArgumentMap& args) {
        if (!m_factories.count(factory)) throw new Exception("No factory registered under specified name!");
        if (m_instances.count(instanceName)) throw new Exception("Instance name must be unique!");
        T* inst = m_factories[factory]->CreateInstance(args);
        if (inst == nullptr) throw new Exception("Instance creation failed!");
        m_instances[instanceName] = inst;
        return inst;
    }
    T* RequestInstance(const std::string& instanceName) {
        auto it = m_instances.find(instanceName);
        if (it == m_instances.end()) throw new Exception("Instance does not exist!");
        return it->second;
    }
    // Deletes the instance. (The reference's ReleaseInstance,
    // FactoryManager.inl:191-203, also throws after a successful release;
    // that is not reproduced.)
    void ReleaseInstance(const std::string& instanceName) {
        auto it = m_instances.find(instanceName);
        if (it == m_instances.end()) throw new Exception("Instance does not exist!");
        delete it->second;
        m_instances.erase(it);
    }
private:
    std::map<std::string, Factory<T>*> m_factories;
    std::map<std::string, T*> m_instances;
};

// The managers the plug-in registers with (Core/System/EngineKernel.h:34-57).
class EngineKernel {
public:
    FactoryManager<IShape>& GetShapeManager() { return m_shapes; }
    FactoryManager<ISpace>& GetSpaceManager() { return m_spaces; }
    FactoryManager<IIntegrator>& GetIntegratorManager() { return m_integrators; }
private:
    // declared so that integrators and spaces are released before shapes
    FactoryManager<IShape> m_shapes;
    FactoryManager<ISpace> m_spaces;
    FactoryManager<IIntegrator> m_integrators;
};

// IPlugIn (Core/System/PlugIn.h:18-38)
class IPlugIn {
public:
    virtual ~IPlugIn() {}
    virtual const std::string& GetName() const = 0;
    virtual const std::string& GetVersion() const = 0;
    virtual void Register(EngineKernel* engine) = 0;
    virtual void Initialise() = 0;
    virtual void Shutdown() = 0;
    virtual void Unregister(EngineKernel* engine) = 0;
};

// Handle of a multi-instance plug-in object (Core/System/LibrarySymbols.h:15).
typedef unsigned long PlugInHandle;

// The device context the plug-in's objects share (created on first use, device 0).
ilm_ctx* GpuContext();

}  // namespace IlluminaGpu

// C symbols the plug-in exports (Core/System/LibrarySymbols.h:15-23,
// Plugin/Dummy/DummyLib.cpp:27-77).
extern "C" {
IlluminaGpu::IPlugIn* GetPlugInSingleton(void);                            // SYMBOL_SINGLE_GET
void DisposePlugInSingleton(void);                                         // SYMBOL_SINGLE_DISPOSE
IlluminaGpu::PlugInHandle CreatePlugInInstance(void);                      // SYMBOL_MULTI_CREATE
IlluminaGpu::IPlugIn* GetPlugInInstance(IlluminaGpu::PlugInHandle handle);  // SYMBOL_MULTI_GET; null if unknown
void DisposePlugInInstance(IlluminaGpu::PlugInHandle handle);              // SYMBOL_MULTI_DISPOSE
void DisposePlugInInstances(void);                                         // SYMBOL_MULTI_DISPOSEALL
// 1 while the plug-in objects hold their shared device context (a
// diagnostic of this plug-in; the reference has no counterpart).
int IlluminaGpuContextHeld(void);
}
