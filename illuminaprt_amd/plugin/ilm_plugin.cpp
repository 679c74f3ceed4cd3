// Implementation of the C++ plug-in layer over the C ABI (ilm_plugin.hpp).
// Built as libilm_plugin.so, linked against libilm_gpu.so.
#include "ilm_plugin.hpp"

#include <cctype>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>

namespace IlluminaGpu {

// ---------------------------------------------------------------- ArgumentMap
// Tokenises on '=' and ';' into name/value pairs (System/ArgumentMap.h:57-79);
// an odd token count is "Invalid argument string!" as there. Surrounding
// whitespace is trimmed.
static std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

ArgumentMap::ArgumentMap(const std::string& text) {
    std::vector<std::string> tok;
    std::string cur;
    for (char c : text) {
        if (c == '=' || c == ';') {
            if (!trim(cur).empty()) tok.push_back(trim(cur));
            cur.clear();
        } else {
            cur += c;
        }
    }
    if (!trim(cur).empty()) tok.push_back(trim(cur));
    if (tok.size() % 2) throw new Exception("Invalid argument string!");
    for (size_t i = 0; i < tok.size(); i += 2) m_args[tok[i]] = tok[i + 1];
}

bool ArgumentMap::GetArgument(const std::string& name, int& v) const {
    auto it = m_args.find(name);
    if (it == m_args.end()) return false;
    std::stringstream ss(it->second);
    int x;
    if (!(ss >> x)) return false;
    v = x;
    return true;
}

bool ArgumentMap::GetArgument(const std::string& name, float& v) const {
    auto it = m_args.find(name);
    if (it == m_args.end()) return false;
    std::stringstream ss(it->second);
    float x;
    if (!(ss >> x)) return false;
    v = x;
    return true;
}

bool ArgumentMap::GetArgument(const std::string& name, std::string& v) const {
    auto it = m_args.find(name);
    if (it == m_args.end()) return false;
    std::stringstream ss(it->second);
    std::string x;
    if (!(ss >> x)) return false;
    v = x;
    return true;
}

bool ArgumentMap::GetArgument(const std::string& name, Vector3& v) const {
    auto it = m_args.find(name);
    if (it == m_args.end()) return false;
    std::stringstream ss(it->second);
    char sep;
    float x, y, z;
    if (!(ss >> sep >> x >> sep >> y >> sep >> z)) return false;
    v = Vector3(x, y, z);
    return true;
}

bool ArgumentMap::GetArgument(const std::string& name, float (&m)[9]) const {
    auto it = m_args.find(name);
    if (it == m_args.end()) return false;
    std::stringstream ss(it->second);
    char sep;
    float e[9];
    if (!(ss >> sep)) return false;
    for (int k = 0; k < 9; ++k) {
        if (!(ss >> e[k])) return false;
        if (!(ss >> sep)) return false;  // ',' between elements, '}' after the last
    }
    for (int k = 0; k < 9; ++k) m[k] = e[k];
    return true;
}

bool ArgumentMap::GetArgument(const std::string& name, std::vector<std::string>& v) const {
    auto it = m_args.find(name);
    if (it == m_args.end()) return true;  // as the reference: an absent list reads as empty
    std::string cur;
    for (char c : it->second + "}") {
        if (c == '{' || c == ',' || c == '}') {
            std::string t = trim(cur);
            if (!t.empty()) v.push_back(t);
            cur.clear();
        } else {
            cur += c;
        }
    }
    return true;
}

bool ArgumentMap::GetArgument(const std::string& name, std::vector<Vector3>& v) const {
    auto it = m_args.find(name);
    if (it == m_args.end()) return true;
    // "{{x, y, z}, ..., {x, y, z}}" (ArgumentMap.h:145-167)
    std::stringstream ss(it->second);
    char sep = 0;
    float x, y, z;
    ss >> sep;
    do {
        if (!(ss >> sep >> x >> sep >> y >> sep >> z >> sep)) return false;
        v.push_back(Vector3(x, y, z));
        if (!(ss >> sep)) return false;
    } while (sep == ',');
    return true;
}

// ---------------------------------------------------------------- context
static std::mutex g_ctx_lock;
static ilm_ctx* g_ctx = nullptr;
// Calls that stage data in an object's buffer and launch work on the shared
// context, one at a time: the reference calls Intersects from several tile
// threads at once, and a context's traversal launches must be ordered.
static std::mutex g_call_lock;

ilm_ctx* GpuContext() {
    std::lock_guard<std::mutex> g(g_ctx_lock);
    if (!g_ctx) Check(ilm_ctx_create(&g_ctx, 0));
    return g_ctx;
}

static void ReleaseGpuContext() {
    std::lock_guard<std::mutex> g(g_ctx_lock);
    if (g_ctx) ilm_ctx_destroy(g_ctx);
    g_ctx = nullptr;
}

// ---------------------------------------------------------------- GpuBVHMesh
GpuBVHMesh::GpuBVHMesh(ilm_ctx* ctx, int maxLeafObjects, int maxTreeDepth, int builder)
    : m_ctx(ctx), m_maxLeaf(maxLeafObjects), m_maxDepth(maxTreeDepth), m_builder(builder) {
    if (!ctx) throw new Exception("GpuBVHMesh needs a device context");
    Check(ilm_alloc(m_ctx, sizeof(ilm_ray) + sizeof(ilm_hit), &m_stage));
}

GpuBVHMesh::~GpuBVHMesh() {
    if (m_mesh) ilm_mesh_destroy(m_mesh);
    if (m_stage) ilm_free(m_ctx, m_stage);
}

void GpuBVHMesh::AddVertex(const Vector3& p) {
    m_pos.push_back(p.X);
    m_pos.push_back(p.Y);
    m_pos.push_back(p.Z);
}

void GpuBVHMesh::AddIndexedTriangle(int v0, int v1, int v2) {
    m_idx.push_back(v0);
    m_idx.push_back(v1);
    m_idx.push_back(v2);
}

bool GpuBVHMesh::Compile() {
    const int64_t nv = (int64_t)m_pos.size() / 3, nt = (int64_t)m_idx.size() / 3;
    if (nt == 0) throw new Exception("GpuBVHMesh::Compile: mesh has no triangles");
    if (m_mesh) {
        ilm_mesh_destroy(m_mesh);
        m_mesh = nullptr;
    }
    Check(ilm_mesh_create(m_ctx, m_pos.data(), nv, m_idx.data(), nt, &m_mesh));
    ilm_bvh_params p;
    std::memset(&p, 0, sizeof p);
    p.builder = m_builder;
    p.width = 4;
    p.max_leaf = m_builder == ILM_BUILDER_REFERENCE ? m_maxLeaf : 0;
    p.max_depth = m_maxDepth;
    Check(ilm_bvh_build(m_mesh, &p));
    return true;
}

static ilm_ray ToRay(const Ray& r) {
    ilm_ray q = {r.Origin.X, r.Origin.Y, r.Origin.Z, r.Min, r.Direction.X, r.Direction.Y, r.Direction.Z, r.Max};
    return q;
}

static void FillSurface(const Ray& r, const ilm_hit& h, DifferentialSurface& s) {
    s.Distance = h.t;
    s.Point = Vector3(r.Origin.X + h.t * r.Direction.X, r.Origin.Y + h.t * r.Direction.Y,
                      r.Origin.Z + h.t * r.Direction.Z);
    s.TriangleIndex = h.prim;
    s.Beta = h.u;
    s.Gamma = h.v;
}

bool GpuBVHMesh::Intersects(const Ray& ray, DifferentialSurface& surface) {
    std::lock_guard<std::mutex> lock(g_call_lock);
    if (!m_mesh) throw new Exception("GpuBVHMesh: Compile() has not been called");
    ilm_ray q = ToRay(ray);
    ilm_hit h;
    ilm_ray* dr = (ilm_ray*)m_stage;
    ilm_hit* dh = (ilm_hit*)((char*)m_stage + sizeof(ilm_ray));
    Check(ilm_upload(m_ctx, dr, &q, sizeof q));
    Check(ilm_trace_closest(m_mesh, dr, 1, dh, nullptr));
    Check(ilm_download(m_ctx, &h, dh, sizeof h));
    if (h.prim < 0) return false;
    FillSurface(ray, h, surface);
    return true;
}

bool GpuBVHMesh::Intersects(const Ray& ray) {
    std::lock_guard<std::mutex> lock(g_call_lock);
    if (!m_mesh) throw new Exception("GpuBVHMesh: Compile() has not been called");
    ilm_ray q = ToRay(ray);
    uint8_t occ = 0;
    ilm_ray* dr = (ilm_ray*)m_stage;
    uint8_t* dh = (uint8_t*)m_stage + sizeof(ilm_ray);
    Check(ilm_upload(m_ctx, dr, &q, sizeof q));
    Check(ilm_trace_any(m_mesh, dr, 1, dh, nullptr));
    Check(ilm_download(m_ctx, &occ, dh, 1));
    return occ != 0;
}

// Host-array batch: one upload, one traversal launch, one download.
void GpuBVHMesh::IntersectBatch(const std::vector<Ray>& rays, std::vector<DifferentialSurface>& out) {
    std::lock_guard<std::mutex> lock(g_call_lock);
    if (!m_mesh) throw new Exception("GpuBVHMesh: Compile() has not been called");
    const int64_t n = (int64_t)rays.size();
    out.assign(rays.size(), DifferentialSurface());
    if (n == 0) return;
    std::vector<ilm_ray> q(n);
    for (int64_t i = 0; i < n; ++i) q[i] = ToRay(rays[i]);
    std::vector<ilm_hit> h(n);
    void* buf = nullptr;
    Check(ilm_alloc(m_ctx, n * (sizeof(ilm_ray) + sizeof(ilm_hit)), &buf));
    ilm_ray* dr = (ilm_ray*)buf;
    ilm_hit* dh = (ilm_hit*)(dr + n);
    int rc = ilm_upload(m_ctx, dr, q.data(), n * sizeof(ilm_ray));
    if (rc == ILM_OK) rc = ilm_trace_closest(m_mesh, dr, n, dh, nullptr);
    if (rc == ILM_OK) rc = ilm_download(m_ctx, h.data(), dh, n * sizeof(ilm_hit));
    ilm_free(m_ctx, buf);
    Check(rc);
    for (int64_t i = 0; i < n; ++i)
        if (h[i].prim >= 0) FillSurface(rays[i], h[i], out[i]);
}

void GpuBVHMesh::OccludedBatch(const std::vector<Ray>& rays, std::vector<uint8_t>& out) {
    std::lock_guard<std::mutex> lock(g_call_lock);
    if (!m_mesh) throw new Exception("GpuBVHMesh: Compile() has not been called");
    const int64_t n = (int64_t)rays.size();
    out.assign(rays.size(), 0);
    if (n == 0) return;
    std::vector<ilm_ray> q(n);
    for (int64_t i = 0; i < n; ++i) q[i] = ToRay(rays[i]);
    void* buf = nullptr;
    Check(ilm_alloc(m_ctx, n * sizeof(ilm_ray) + n, &buf));
    ilm_ray* dr = (ilm_ray*)buf;
    uint8_t* dh = (uint8_t*)(dr + n);
    int rc = ilm_upload(m_ctx, dr, q.data(), n * sizeof(ilm_ray));
    if (rc == ILM_OK) rc = ilm_trace_any(m_mesh, dr, n, dh, nullptr);
    if (rc == ILM_OK) rc = ilm_download(m_ctx, out.data(), dh, n);
    ilm_free(m_ctx, buf);
    Check(rc);
}

// ---------------------------------------------------------------- GpuSpace
GpuSpace::GpuSpace(ilm_ctx* ctx) : m_ctx(ctx) {
    if (!ctx) throw new Exception("GpuSpace needs a device context");
    Check(ilm_alloc(m_ctx, sizeof(ilm_ray) + sizeof(ilm_hit) + 2 * sizeof(int32_t), &m_stage));
}

GpuSpace::~GpuSpace() {
    if (m_space) ilm_space_destroy(m_space);
    if (m_stage) ilm_free(m_ctx, m_stage);
}

bool GpuSpace::Build() {
    if (PrimitiveList.empty()) throw new Exception("GpuSpace::Build: no primitives");
    if (m_space) {
        ilm_space_destroy(m_space);
        m_space = nullptr;
        m_mesh = nullptr;
    }
    Check(ilm_space_create(m_ctx, &m_space));
    for (const Primitive& p : PrimitiveList) {
        if (!p.Shape) throw new Exception("GpuSpace::Build: primitive without a shape");
        const std::vector<float>& pos = p.Shape->Positions();
        const std::vector<int32_t>& idx = p.Shape->Indices();
        Check(ilm_space_add(m_space, pos.data(), (int64_t)pos.size() / 3, idx.data(), (int64_t)idx.size() / 3,
                            nullptr, p.Rotation, p.Scaling, p.Translation, p.Material, p.Light, nullptr));
    }
    ilm_bvh_params bp;
    std::memset(&bp, 0, sizeof bp);
    bp.builder = ILM_BUILDER_SAH;
    Check(ilm_space_build(m_space, &bp));
    Check(ilm_space_mesh(m_space, &m_mesh));
    return true;
}

bool GpuSpace::Intersects(const Ray& ray, Intersection& intersection) {
    std::lock_guard<std::mutex> lock(g_call_lock);
    if (!m_mesh) throw new Exception("GpuSpace: Build() has not been called");
    ilm_ray q = ToRay(ray);
    ilm_ray* dr = (ilm_ray*)m_stage;
    ilm_hit* dh = (ilm_hit*)(dr + 1);
    int32_t* dp = (int32_t*)(dh + 1);
    Check(ilm_upload(m_ctx, dr, &q, sizeof q));
    Check(ilm_trace_closest(m_mesh, dr, 1, dh, nullptr));
    Check(ilm_space_resolve(m_space, dh, 1, dp, dp + 1, nullptr));
    struct {
        ilm_hit h;
        int32_t prim, local;
    } r;
    Check(ilm_download(m_ctx, &r, dh, sizeof r));
    if (r.h.prim < 0) return false;
    FillSurface(ray, r.h, intersection.Surface);
    intersection.Surface.TriangleIndex = r.local;
    intersection.PrimitiveIndex = r.prim;
    return true;
}

bool GpuSpace::Intersects(const Ray& ray) {
    std::lock_guard<std::mutex> lock(g_call_lock);
    if (!m_mesh) throw new Exception("GpuSpace: Build() has not been called");
    ilm_ray q = ToRay(ray);
    uint8_t occ = 0;
    ilm_ray* dr = (ilm_ray*)m_stage;
    uint8_t* dh = (uint8_t*)(dr + 1);
    Check(ilm_upload(m_ctx, dr, &q, sizeof q));
    Check(ilm_trace_any(m_mesh, dr, 1, dh, nullptr));
    Check(ilm_download(m_ctx, &occ, dh, 1));
    return occ != 0;
}

void GpuSpace::IntersectBatch(const std::vector<Ray>& rays, std::vector<Intersection>& out) {
    std::lock_guard<std::mutex> lock(g_call_lock);
    if (!m_mesh) throw new Exception("GpuSpace: Build() has not been called");
    const int64_t n = (int64_t)rays.size();
    out.assign(rays.size(), Intersection());
    if (n == 0) return;
    std::vector<ilm_ray> q(n);
    for (int64_t i = 0; i < n; ++i) q[i] = ToRay(rays[i]);
    std::vector<ilm_hit> h(n);
    std::vector<int32_t> pl(2 * n);
    void* buf = nullptr;
    Check(ilm_alloc(m_ctx, n * (sizeof(ilm_ray) + sizeof(ilm_hit) + 2 * sizeof(int32_t)), &buf));
    ilm_ray* dr = (ilm_ray*)buf;
    ilm_hit* dh = (ilm_hit*)(dr + n);
    int32_t* dp = (int32_t*)(dh + n);
    int rc = ilm_upload(m_ctx, dr, q.data(), n * sizeof(ilm_ray));
    if (rc == ILM_OK) rc = ilm_trace_closest(m_mesh, dr, n, dh, nullptr);
    if (rc == ILM_OK) rc = ilm_space_resolve(m_space, dh, n, dp, dp + n, nullptr);
    if (rc == ILM_OK) rc = ilm_download(m_ctx, h.data(), dh, n * sizeof(ilm_hit));
    if (rc == ILM_OK) rc = ilm_download(m_ctx, pl.data(), dp, 2 * n * sizeof(int32_t));
    ilm_free(m_ctx, buf);
    Check(rc);
    for (int64_t i = 0; i < n; ++i) {
        if (h[i].prim < 0) continue;
        FillSurface(rays[i], h[i], out[i].Surface);
        out[i].Surface.TriangleIndex = pl[n + i];
        out[i].PrimitiveIndex = pl[i];
    }
}

// ---------------------------------------------------------------- GpuPathTracing
GpuPathTracing::GpuPathTracing(ilm_ctx* ctx, int rayDepth, float epsilon)
    : m_ctx(ctx), m_rayDepth(rayDepth), m_epsilon(epsilon) {
    if (!ctx) throw new Exception("GpuPathTracing needs a device context");
    if (rayDepth < 1) throw new Exception("GpuPathTracing: RayDepth must be at least 1");
}

GpuPathTracing::~GpuPathTracing() {
    if (m_scene) ilm_scene_destroy(m_scene);
}

void GpuPathTracing::Initialise(GpuBVHMesh* mesh, const std::vector<float>& vertexNormals,
                                const std::vector<int32_t>& triMaterial, const std::vector<int32_t>& triLight,
                                const std::vector<float>& materials, const std::vector<float>& lights) {
    if (!mesh || !mesh->Handle()) throw new Exception("GpuPathTracing::Initialise: mesh is not compiled");
    const size_t nt = mesh->Indices().size() / 3;
    if (vertexNormals.size() != mesh->Positions().size() || triMaterial.size() != nt || triLight.size() != nt ||
        materials.size() % 8 || lights.size() % 4)
        throw new Exception("GpuPathTracing::Initialise: inconsistent scene arrays");
    if (m_scene) {
        ilm_scene_destroy(m_scene);
        m_scene = nullptr;
    }
    Check(ilm_scene_create(m_ctx, mesh->Handle(), vertexNormals.data(), triMaterial.data(), triLight.data(),
                           materials.data(), (int)(materials.size() / 8), lights.data(), (int)(lights.size() / 4),
                           &m_scene));
}

void GpuPathTracing::RenderRegion(const CameraArgs& cam, int width, int height, int spp, uint32_t seed,
                                  std::vector<float>& image) {
    std::lock_guard<std::mutex> lock(g_call_lock);
    if (!m_scene) throw new Exception("GpuPathTracing: Initialise() has not been called");
    if (width <= 0 || height <= 0 || spp <= 0) throw new Exception("GpuPathTracing: bad image size or spp");
    ilm_render_params p;
    std::memset(&p, 0, sizeof p);
    const Vector3* v[3] = {&cam.Position, &cam.Target, &cam.Up};
    float* dst[3] = {p.cam.pos, p.cam.target, p.cam.up};
    for (int k = 0; k < 3; ++k) {
        dst[k][0] = v[k]->X;
        dst[k][1] = v[k]->Y;
        dst[k][2] = v[k]->Z;
    }
    p.cam.fov_deg = cam.Fov;
    p.cam.aspect = cam.Aspect;
    p.width = width;
    p.height = height;
    p.spp = spp;
    p.max_depth = m_rayDepth;
    p.rr_depth = 4;  // PathIntegrator::Radiance rolls the roulette when depth > 3
    p.eps = m_epsilon;
    p.seed = seed;
    const size_t npx = (size_t)width * height;
    void* accum = nullptr;
    Check(ilm_alloc(m_ctx, npx * 3 * sizeof(double), &accum));
    int rc = ilm_memset(m_ctx, accum, 0, npx * 3 * sizeof(double));
    if (rc == ILM_OK) rc = ilm_render(m_scene, &p, (double*)accum, nullptr);
    std::vector<double> sum(npx * 3);
    if (rc == ILM_OK) rc = ilm_download(m_ctx, sum.data(), accum, sum.size() * sizeof(double));
    ilm_free(m_ctx, accum);
    Check(rc);
    image.resize(npx * 3);
    for (size_t i = 0; i < sum.size(); ++i) image[i] = (float)(sum[i] / spp);
}

// ---------------------------------------------------------------- factories
// Shape arguments as BVHMeshShapeFactory reads them (Shape/ShapeFactories.h:304-322),
// plus Builder = SAH | Reference (default SAH).
class GpuBVHMeshFactory : public Factory<IShape> {
public:
    IShape* CreateInstance(ArgumentMap& args) override {
        int maxDepth = 35, maxObjects = 10;
        std::string builder = "SAH";
        args.GetArgument("MaximumTreeDepth", maxDepth);
        args.GetArgument("MaximumLeafObjects", maxObjects);
        args.GetArgument("Builder", builder);
        int b;
        if (builder == "SAH")
            b = ILM_BUILDER_SAH;
        else if (builder == "Reference")
            b = ILM_BUILDER_REFERENCE;
        else
            throw new Exception("GpuBVHMesh: Builder must be SAH or Reference");
        return new GpuBVHMesh(GpuContext(), maxObjects, maxDepth, b);
    }
};

// Registered as "GpuSpace" in the place of the "Basic" space factory
// (Space/SpaceFactories.h); it takes no arguments.
class GpuSpaceFactory : public Factory<ISpace> {
public:
    ISpace* CreateInstance(ArgumentMap&) override { return new GpuSpace(GpuContext()); }
};

// Integrator arguments as PathIntegratorFactory reads them
// (Integrator/IntegratorFactories.h:41-57). One shadow ray per light per
// vertex is what the GPU integrator traces; another ShadowRays value throws.
class GpuPathTracingFactory : public Factory<IIntegrator> {
public:
    IIntegrator* CreateInstance(ArgumentMap& args) override {
        int rayDepth = 6, shadowRays = 1;
        float epsilon = 1e-4f;
        args.GetArgument("RayDepth", rayDepth);
        args.GetArgument("ShadowRays", shadowRays);
        args.GetArgument("Epsilon", epsilon);
        if (shadowRays != 1) throw new Exception("GpuPathTracing: only ShadowRays = 1 is supported");
        return new GpuPathTracing(GpuContext(), rayDepth, epsilon);
    }
};

// ---------------------------------------------------------------- plug-in
// Registers and unregisters its factories the way the sample plug-in does
// (Plugin/Dummy/DummyPlugIn.h:54-111).
class GpuAccelPlugIn : public IPlugIn {
public:
    const std::string& GetName() const override { return m_name; }
    const std::string& GetVersion() const override { return m_version; }
    void Register(EngineKernel* engine) override {
        engine->GetShapeManager().RegisterFactory("GpuBVHMesh", &m_shape);
        engine->GetSpaceManager().RegisterFactory("GpuSpace", &m_space);
        engine->GetIntegratorManager().RegisterFactory("GpuPathTracing", &m_integrator);
    }
    void Initialise() override { GpuContext(); }
    void Shutdown() override {}
    void Unregister(EngineKernel* engine) override {
        engine->GetShapeManager().UnregisterFactory("GpuBVHMesh");
        engine->GetSpaceManager().UnregisterFactory("GpuSpace");
        engine->GetIntegratorManager().UnregisterFactory("GpuPathTracing");
    }
private:
    std::string m_name = "GpuAccel";
    std::string m_version = "1.0.0";
    GpuBVHMeshFactory m_shape;
    GpuSpaceFactory m_space;
    GpuPathTracingFactory m_integrator;
};

}  // namespace IlluminaGpu

// Plug-in objects: the singleton (GetPlugInSingleton / DisposePlugInSingleton)
// and the multi-instance objects (CreatePlugInInstance ... DisposePlugInInstances,
// keyed by handle). The device context they share is released once the last
// of them is disposed; the engine releases the plug-in's shapes, spaces and
// integrators before it disposes the plug-in.
static std::mutex g_plugin_lock;
static IlluminaGpu::GpuAccelPlugIn* g_plugin = nullptr;
static std::map<IlluminaGpu::PlugInHandle, IlluminaGpu::GpuAccelPlugIn*> g_instances;
static IlluminaGpu::PlugInHandle g_next_handle = 1;

// Caller holds g_plugin_lock.
static void ReleaseContextIfUnused() {
    if (!g_plugin && g_instances.empty()) IlluminaGpu::ReleaseGpuContext();
}

extern "C" IlluminaGpu::IPlugIn* GetPlugInSingleton(void) {
    std::lock_guard<std::mutex> g(g_plugin_lock);
    if (!g_plugin) g_plugin = new IlluminaGpu::GpuAccelPlugIn();
    return g_plugin;
}

extern "C" void DisposePlugInSingleton(void) {
    std::lock_guard<std::mutex> g(g_plugin_lock);
    delete g_plugin;
    g_plugin = nullptr;
    ReleaseContextIfUnused();
}

extern "C" IlluminaGpu::PlugInHandle CreatePlugInInstance(void) {
    std::lock_guard<std::mutex> g(g_plugin_lock);
    const IlluminaGpu::PlugInHandle h = g_next_handle++;
    g_instances[h] = new IlluminaGpu::GpuAccelPlugIn();
    return h;
}

extern "C" IlluminaGpu::IPlugIn* GetPlugInInstance(IlluminaGpu::PlugInHandle handle) {
    std::lock_guard<std::mutex> g(g_plugin_lock);
    auto it = g_instances.find(handle);
    return it == g_instances.end() ? nullptr : it->second;
}

extern "C" void DisposePlugInInstance(IlluminaGpu::PlugInHandle handle) {
    std::lock_guard<std::mutex> g(g_plugin_lock);
    auto it = g_instances.find(handle);
    if (it == g_instances.end()) return;
    delete it->second;
    g_instances.erase(it);
    ReleaseContextIfUnused();
}

extern "C" void DisposePlugInInstances(void) {
    std::lock_guard<std::mutex> g(g_plugin_lock);
    for (auto& kv : g_instances) delete kv.second;
    g_instances.clear();
    ReleaseContextIfUnused();
}

extern "C" int IlluminaGpuContextHeld(void) {
    std::lock_guard<std::mutex> g(IlluminaGpu::g_ctx_lock);
    return IlluminaGpu::g_ctx != nullptr;
}
