// A small host program that drives the GpuAccel plug-in the way the
// reference's sandbox does: it loads the plug-in library by path and
// resolves GetPlugInSingleton (System/PlugInManager.cpp:40-96 with
// System/LibrarySymbols.h:15-23), registers its factories with an
// EngineKernel, creates a shape and an integrator through the factory
// managers from argument strings (Scene/EnvironmentLoader.cpp:567-647), and
// runs them.
//
//   plugin_sandbox <plugin.so> host
//       host-side checks (argument parsing, factory errors, registration);
//       needs no GPU. Prints "host ok" on success.
//   plugin_sandbox <plugin.so> multi
//       resolves all six plug-in symbols through the loader's types and
//       checks singleton / multi-instance lifetimes and the release of the
//       shared device context (without a device: the bookkeeping only).
//       Prints "multi ok" on success.
//   plugin_sandbox <plugin.so> trace <in.bin> <out.bin> "<shape args>"
//       in:  int64 nv, nt, n, nper; float pos[3nv]; int32 idx[3nt]; float rays[8n]
//       out: batched closest hits (n x {t, prim, u, v}), batched occlusion
//            (n bytes), per-ray closest hits and occlusion for the first
//            nper rays (IShape::Intersects one ray at a time).
//   plugin_sandbox <plugin.so> space <in.bin> <out.bin> ""
//       in:  int64 nprims, n, nper; per primitive: int64 nv, nt; float rot[9],
//            scl[3], trn[3]; float pos[3nv]; int32 idx[3nt]; then float rays[8n]
//       out: n x {t, primitive, triangle} (batched), then the same for the
//            first nper rays through the per-ray ISpace::Intersects, then
//            nper occlusion bytes.
//   plugin_sandbox <plugin.so> render <in.bin> <out.bin> "<integrator args>"
//       in:  int64 nv, nt, nmat, nlight; int32 width, height, spp, seed;
//            float camera[11] (pos, target, up, fov, aspect); float pos[3nv];
//            int32 idx[3nt]; float nrm[3nv]; int32 tri_mat[nt], tri_light[nt];
//            float mats[8 nmat], lights[4 nlight]
//       out: float image[width * height * 3] (mean radiance per pixel)
//   plugin_sandbox <plugin.so> ilm-load <script.ilm> <out.bin>
//       loads a scene script on the host (no GPU) and writes the resolved job:
//       int64 nv, nt, nmat, nlight; int32 width, height, spp, ray depth,
//       builder; float epsilon; float camera[11]; float pos[3nv]; int32
//       idx[3nt]; float nrm[3nv]; int32 tri_mat[nt], tri_light[nt]; float
//       mats[8 nmat], lights[4 nlight]; int64 nprims; int64
//       first_triangle[nprims + 1]. Prints the loader's warnings.
//   plugin_sandbox <plugin.so> ilm-render <script.ilm> <seed> <out.bin>
//       loads the script, renders it on the GPU, writes the image device's
//       file (PPM, or PFM with Format = PFM) and the mean radiance in the
//       device's pixel order to out.bin.
//
// Exit status 0 on success; a plug-in exception prints its message and exits 3.
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "ilm_environment.hpp"
#include "ilm_plugin.hpp"

using namespace IlluminaGpu;

// The loader's symbol types (Core/System/LibrarySymbols.h:17-23).
typedef IPlugIn* (*SYMBOL_SINGLE_GET)(void);
typedef void (*SYMBOL_SINGLE_DISPOSE)(void);
typedef PlugInHandle (*SYMBOL_MULTI_CREATE)(void);
typedef IPlugIn* (*SYMBOL_MULTI_GET)(PlugInHandle);
typedef void (*SYMBOL_MULTI_DISPOSE)(PlugInHandle);
typedef void (*SYMBOL_MULTI_DISPOSEALL)(void);
typedef int (*ContextHeldFn)(void);

template <class T>
static void read_into(FILE* f, std::vector<T>& v, size_t n) {
    v.resize(n);
    if (n && fread(v.data(), sizeof(T), n, f) != n) throw new Exception("short input file");
}

template <class T>
static T read_one(FILE* f) {
    T x;
    if (fread(&x, sizeof x, 1, f) != 1) throw new Exception("short input file");
    return x;
}

static bool expect_throw(void (*fn)(EngineKernel&), EngineKernel& k, const char* what) {
    try {
        fn(k);
    } catch (Exception* e) {
        delete e;
        return true;
    }
    fprintf(stderr, "expected an exception: %s\n", what);
    return false;
}

static int run_host(IPlugIn* plugin) {
    bool ok = true;
    // Argument parsing as the loader feeds it.
    ArgumentMap a("MaximumTreeDepth = 20; MaximumLeafObjects=4; Builder = Reference; Epsilon = 0.001;");
    int depth = 0, leaf = 0;
    float eps = 0;
    std::string builder;
    ok &= a.GetArgument("MaximumTreeDepth", depth) && depth == 20;
    ok &= a.GetArgument("MaximumLeafObjects", leaf) && leaf == 4;
    ok &= a.GetArgument("Builder", builder) && builder == "Reference";
    ok &= a.GetArgument("Epsilon", eps) && eps == 0.001f;
    ok &= !a.GetArgument("Missing", depth) && depth == 20;
    try {
        ArgumentMap bad("A = 1; B");
        ok = false;
        fprintf(stderr, "odd token count accepted\n");
    } catch (Exception* e) {
        delete e;
    }

    EngineKernel kernel;
    plugin->Register(&kernel);
    ok &= kernel.GetShapeManager().QueryFactory("GpuBVHMesh");
    ok &= kernel.GetIntegratorManager().QueryFactory("GpuPathTracing");
    ok &= kernel.GetSpaceManager().QueryFactory("GpuSpace");
    ok &= expect_throw([](EngineKernel& k) { k.GetShapeManager().RegisterFactory("GpuBVHMesh", nullptr); }, kernel,
                       "duplicate factory");
    ok &= expect_throw(
        [](EngineKernel& k) {
            ArgumentMap m;
            k.GetShapeManager().CreateInstance("NoSuchFactory", "x", m);
        },
        kernel, "unknown factory");
    ok &= expect_throw([](EngineKernel& k) { k.GetShapeManager().RequestInstance("nothing"); }, kernel,
                       "unknown instance");
    int ndev = 0;
    ilm_device_count(&ndev);
    if (ndev == 0) {
        // No device: creating a GPU shape must fail loudly, never fall back.
        ok &= expect_throw(
            [](EngineKernel& k) {
                ArgumentMap m("Builder = SAH");
                k.GetShapeManager().CreateInstance("GpuBVHMesh", "mesh", m);
            },
            kernel, "shape without a device");
    }
    ok &= expect_throw(
        [](EngineKernel& k) {
            ArgumentMap m("ShadowRays = 4");
            k.GetIntegratorManager().CreateInstance("GpuPathTracing", "pt", m);
        },
        kernel, "unsupported ShadowRays");
    plugin->Unregister(&kernel);
    ok &= !kernel.GetShapeManager().QueryFactory("GpuBVHMesh");
    printf(ok ? "host ok\n" : "host FAILED\n");
    return ok ? 0 : 1;
}

// Resolves all six plug-in symbols through the loader's types and checks the
// lifetime rules: handles name distinct objects until disposed, and the shared
// device context is released (with a device) once the singleton and every
// instance are gone. The singleton is disposed with no argument, as
// LibraryManager::DisposePlugInSingleton calls it (LibraryManager.cpp:80-90).
static int run_multi(void* lib) {
    auto get = (SYMBOL_SINGLE_GET)dlsym(lib, "GetPlugInSingleton");
    auto dispose = (SYMBOL_SINGLE_DISPOSE)dlsym(lib, "DisposePlugInSingleton");
    auto create = (SYMBOL_MULTI_CREATE)dlsym(lib, "CreatePlugInInstance");
    auto get_inst = (SYMBOL_MULTI_GET)dlsym(lib, "GetPlugInInstance");
    auto dispose_inst = (SYMBOL_MULTI_DISPOSE)dlsym(lib, "DisposePlugInInstance");
    auto dispose_all = (SYMBOL_MULTI_DISPOSEALL)dlsym(lib, "DisposePlugInInstances");
    auto held = (ContextHeldFn)dlsym(lib, "IlluminaGpuContextHeld");
    if (!get || !dispose || !create || !get_inst || !dispose_inst || !dispose_all || !held) {
        fprintf(stderr, "multi: a plug-in symbol is missing\n");
        return 1;
    }
    int ndev = 0;
    ilm_device_count(&ndev);
    bool ok = true;
    auto check = [&](bool c, const char* what) {
        if (!c) fprintf(stderr, "multi: %s\n", what);
        ok &= c;
    };
    IPlugIn* s = get();
    check(s && s == get() && s->GetName() == "GpuAccel", "singleton");
    if (ndev) {
        s->Initialise();
        check(held() == 1, "singleton holds the context after Initialise");
    }
    dispose();
    check(held() == 0, "no-argument DisposePlugInSingleton releases the context");
    const PlugInHandle h1 = create(), h2 = create();
    IPlugIn* p1 = get_inst(h1);
    IPlugIn* p2 = get_inst(h2);
    check(h1 != h2 && p1 && p2 && p1 != p2, "instances are distinct");
    check(get_inst(h1) == p1 && get_inst(h1 + h2 + 100) == nullptr, "handles name their instance");
    EngineKernel kernel;
    p1->Register(&kernel);
    check(kernel.GetShapeManager().QueryFactory("GpuBVHMesh"), "an instance registers its factories");
    p1->Unregister(&kernel);
    if (ndev) {
        p1->Initialise();
        check(held() == 1, "instance holds the context");
    }
    dispose_inst(h1);
    check(get_inst(h1) == nullptr && get_inst(h2) == p2, "DisposePlugInInstance disposes one instance");
    if (ndev) check(held() == 1, "the context stays while an instance remains");
    dispose_inst(h1);  // an unknown handle is ignored
    dispose_all();
    check(get_inst(h2) == nullptr, "DisposePlugInInstances disposes every instance");
    check(held() == 0, "the context is released with the last instance");
    IPlugIn* s2 = get();
    check(s2 && s2->GetName() == "GpuAccel", "a new singleton after disposal");
    dispose();
    printf(ok ? "multi ok\n" : "multi FAILED\n");
    return ok ? 0 : 1;
}

static int run_trace(IPlugIn* plugin, const char* in, const char* out, const char* args) {
    FILE* f = fopen(in, "rb");
    if (!f) throw new Exception(std::string("cannot open ") + in);
    const int64_t nv = read_one<int64_t>(f), nt = read_one<int64_t>(f), n = read_one<int64_t>(f),
                  nper = read_one<int64_t>(f);
    std::vector<float> pos, rays;
    std::vector<int32_t> idx;
    read_into(f, pos, (size_t)nv * 3);
    read_into(f, idx, (size_t)nt * 3);
    read_into(f, rays, (size_t)n * 8);
    fclose(f);
    if (nper > n) throw new Exception("nper > n");

    EngineKernel kernel;
    plugin->Register(&kernel);
    plugin->Initialise();
    ArgumentMap am(args);
    GpuBVHMesh* mesh = dynamic_cast<GpuBVHMesh*>(kernel.GetShapeManager().CreateInstance("GpuBVHMesh", "mesh", am));
    if (!mesh) throw new Exception("factory returned another shape type");
    for (int64_t v = 0; v < nv; ++v) mesh->AddVertex(Vector3(pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]));
    for (int64_t t = 0; t < nt; ++t) mesh->AddIndexedTriangle(idx[3 * t], idx[3 * t + 1], idx[3 * t + 2]);
    mesh->Compile();

    std::vector<Ray> rs(n);
    for (int64_t i = 0; i < n; ++i) {
        const float* r = &rays[8 * i];
        rs[i].Origin = Vector3(r[0], r[1], r[2]);
        rs[i].Min = r[3];
        rs[i].Direction = Vector3(r[4], r[5], r[6]);
        rs[i].Max = r[7];
    }
    std::vector<DifferentialSurface> hits;
    std::vector<uint8_t> occ;
    mesh->IntersectBatch(rs, hits);
    mesh->OccludedBatch(rs, occ);

    FILE* o = fopen(out, "wb");
    if (!o) throw new Exception(std::string("cannot open ") + out);
    auto put_hit = [&](const DifferentialSurface& s) {
        float rec[4] = {s.Distance, 0.0f, s.Beta, s.Gamma};
        int32_t prim = s.TriangleIndex;
        memcpy(&rec[1], &prim, 4);
        fwrite(rec, sizeof rec, 1, o);
    };
    for (int64_t i = 0; i < n; ++i) put_hit(hits[i]);
    fwrite(occ.data(), 1, (size_t)n, o);
    for (int64_t i = 0; i < nper; ++i) {
        DifferentialSurface s;
        if (!mesh->Intersects(rs[i], s)) s = DifferentialSurface();
        put_hit(s);
    }
    for (int64_t i = 0; i < nper; ++i) {
        uint8_t b = mesh->Intersects(rs[i]) ? 1 : 0;
        fwrite(&b, 1, 1, o);
    }
    fclose(o);
    kernel.GetShapeManager().ReleaseInstance("mesh");
    plugin->Shutdown();
    plugin->Unregister(&kernel);
    printf("trace ok: %lld rays, %lld per-ray\n", (long long)n, (long long)nper);
    return 0;
}

static int run_space(IPlugIn* plugin, const char* in, const char* out) {
    FILE* f = fopen(in, "rb");
    if (!f) throw new Exception(std::string("cannot open ") + in);
    const int64_t np = read_one<int64_t>(f), n = read_one<int64_t>(f), nper = read_one<int64_t>(f);
    if (nper > n) throw new Exception("nper > n");
    EngineKernel kernel;
    plugin->Register(&kernel);
    plugin->Initialise();
    ArgumentMap none;
    GpuSpace* space = dynamic_cast<GpuSpace*>(kernel.GetSpaceManager().CreateInstance("GpuSpace", "space", none));
    if (!space) throw new Exception("factory returned another space type");
    for (int64_t k = 0; k < np; ++k) {
        const int64_t nv = read_one<int64_t>(f), nt = read_one<int64_t>(f);
        std::vector<float> xf, pos;
        std::vector<int32_t> idx;
        read_into(f, xf, 15);
        read_into(f, pos, (size_t)nv * 3);
        read_into(f, idx, (size_t)nt * 3);
        ArgumentMap shapeArgs("Builder = SAH");
        GpuBVHMesh* mesh = dynamic_cast<GpuBVHMesh*>(
            kernel.GetShapeManager().CreateInstance("GpuBVHMesh", "shape" + std::to_string(k), shapeArgs));
        for (int64_t v = 0; v < nv; ++v) mesh->AddVertex(Vector3(pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]));
        for (int64_t t = 0; t < nt; ++t) mesh->AddIndexedTriangle(idx[3 * t], idx[3 * t + 1], idx[3 * t + 2]);
        Primitive p;
        p.Shape = mesh;
        memcpy(p.Rotation, &xf[0], sizeof p.Rotation);
        memcpy(p.Scaling, &xf[9], sizeof p.Scaling);
        memcpy(p.Translation, &xf[12], sizeof p.Translation);
        space->PrimitiveList.push_back(p);
    }
    std::vector<float> rays;
    read_into(f, rays, (size_t)n * 8);
    fclose(f);
    space->Initialise();
    space->Build();
    std::vector<Ray> rs(n);
    for (int64_t i = 0; i < n; ++i) {
        const float* r = &rays[8 * i];
        rs[i].Origin = Vector3(r[0], r[1], r[2]);
        rs[i].Min = r[3];
        rs[i].Direction = Vector3(r[4], r[5], r[6]);
        rs[i].Max = r[7];
    }
    std::vector<Intersection> hits;
    space->IntersectBatch(rs, hits);
    FILE* o = fopen(out, "wb");
    if (!o) throw new Exception(std::string("cannot open ") + out);
    auto put = [&](const Intersection& x) {
        float t = x.Surface.Distance;
        int32_t rec[2] = {x.PrimitiveIndex, x.Surface.TriangleIndex};
        fwrite(&t, 4, 1, o);
        fwrite(rec, 4, 2, o);
    };
    for (int64_t i = 0; i < n; ++i) put(hits[i]);
    for (int64_t i = 0; i < nper; ++i) {
        Intersection x;
        if (!space->Intersects(rs[i], x)) x = Intersection();
        put(x);
    }
    for (int64_t i = 0; i < nper; ++i) {
        uint8_t b = space->Intersects(rs[i]) ? 1 : 0;
        fwrite(&b, 1, 1, o);
    }
    fclose(o);
    space->Shutdown();
    kernel.GetSpaceManager().ReleaseInstance("space");
    plugin->Shutdown();
    plugin->Unregister(&kernel);
    printf("space ok: %lld primitives, %lld rays\n", (long long)np, (long long)n);
    return 0;
}

static int run_render(IPlugIn* plugin, const char* in, const char* out, const char* args) {
    FILE* f = fopen(in, "rb");
    if (!f) throw new Exception(std::string("cannot open ") + in);
    const int64_t nv = read_one<int64_t>(f), nt = read_one<int64_t>(f), nmat = read_one<int64_t>(f),
                  nlight = read_one<int64_t>(f);
    const int32_t W = read_one<int32_t>(f), H = read_one<int32_t>(f), spp = read_one<int32_t>(f),
                  seed = read_one<int32_t>(f);
    std::vector<float> cam, pos, nrm, mats, lights;
    std::vector<int32_t> idx, tmat, tlight;
    read_into(f, cam, 11);
    read_into(f, pos, (size_t)nv * 3);
    read_into(f, idx, (size_t)nt * 3);
    read_into(f, nrm, (size_t)nv * 3);
    read_into(f, tmat, (size_t)nt);
    read_into(f, tlight, (size_t)nt);
    read_into(f, mats, (size_t)nmat * 8);
    read_into(f, lights, (size_t)nlight * 4);
    fclose(f);

    EngineKernel kernel;
    plugin->Register(&kernel);
    plugin->Initialise();
    ArgumentMap shapeArgs("Builder = SAH");
    GpuBVHMesh* mesh =
        dynamic_cast<GpuBVHMesh*>(kernel.GetShapeManager().CreateInstance("GpuBVHMesh", "scene", shapeArgs));
    for (int64_t v = 0; v < nv; ++v) mesh->AddVertex(Vector3(pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]));
    for (int64_t t = 0; t < nt; ++t) mesh->AddIndexedTriangle(idx[3 * t], idx[3 * t + 1], idx[3 * t + 2]);
    mesh->Compile();
    ArgumentMap am(args);
    GpuPathTracing* pt =
        dynamic_cast<GpuPathTracing*>(kernel.GetIntegratorManager().CreateInstance("GpuPathTracing", "pt", am));
    if (!mesh || !pt) throw new Exception("factory returned another type");
    pt->Initialise(mesh, nrm, tmat, tlight, mats, lights);
    CameraArgs c;
    c.Position = Vector3(cam[0], cam[1], cam[2]);
    c.Target = Vector3(cam[3], cam[4], cam[5]);
    c.Up = Vector3(cam[6], cam[7], cam[8]);
    c.Fov = cam[9];
    c.Aspect = cam[10];
    std::vector<float> image;
    pt->RenderRegion(c, W, H, spp, (uint32_t)seed, image);
    FILE* o = fopen(out, "wb");
    if (!o) throw new Exception(std::string("cannot open ") + out);
    fwrite(image.data(), sizeof(float), image.size(), o);
    fclose(o);
    // The integrator refers to the mesh's device arrays: release it first.
    kernel.GetIntegratorManager().ReleaseInstance("pt");
    kernel.GetShapeManager().ReleaseInstance("scene");
    plugin->Shutdown();
    plugin->Unregister(&kernel);
    printf("render ok: %dx%d, %d spp\n", W, H, spp);
    return 0;
}

template <class T>
static void write_vec(FILE* o, const std::vector<T>& v) {
    if (!v.empty()) fwrite(v.data(), sizeof(T), v.size(), o);
}

static int run_ilm_load(const char* script, const char* out) {
    EnvironmentLoader loader;
    if (!loader.Load(std::string(script))) throw new Exception(std::string("cannot open ") + script);
    for (const std::string& w : loader.Warnings()) printf("warning: %s\n", w.c_str());
    const RenderJob& j = loader.Job();
    FILE* o = fopen(out, "wb");
    if (!o) throw new Exception(std::string("cannot open ") + out);
    const int64_t head[4] = {(int64_t)j.Positions.size() / 3, (int64_t)j.Indices.size() / 3,
                             (int64_t)j.Materials.size() / 8, (int64_t)j.Lights.size() / 4};
    const int32_t ints[5] = {j.Width, j.Height, j.Samples, j.RayDepth, j.Builder};
    const ilm_camera& c = j.Camera;
    const float cam[11] = {c.pos[0], c.pos[1], c.pos[2], c.target[0], c.target[1], c.target[2],
                           c.up[0],  c.up[1],  c.up[2],  c.fov_deg,   c.aspect};
    fwrite(head, sizeof head, 1, o);
    fwrite(ints, sizeof ints, 1, o);
    fwrite(&j.Epsilon, sizeof(float), 1, o);
    fwrite(cam, sizeof cam, 1, o);
    write_vec(o, j.Positions);
    write_vec(o, j.Indices);
    write_vec(o, j.Normals);
    write_vec(o, j.TriMaterial);
    write_vec(o, j.TriLight);
    write_vec(o, j.Materials);
    write_vec(o, j.Lights);
    const int64_t np = (int64_t)j.PrimitiveIds.size();
    fwrite(&np, sizeof np, 1, o);
    write_vec(o, j.PrimitiveFirstTriangle);
    fclose(o);
    printf("ilm-load ok: %lld primitives, %lld triangles, %lld lights\n", (long long)np, (long long)head[1],
           (long long)head[3]);
    return 0;
}

static int run_ilm_render(const char* script, const char* seed, const char* out) {
    EnvironmentLoader loader;
    if (!loader.Load(std::string(script))) throw new Exception(std::string("cannot open ") + script);
    for (const std::string& w : loader.Warnings()) printf("warning: %s\n", w.c_str());
    std::vector<float> image;
    loader.Job().Render((uint32_t)strtoul(seed, nullptr, 10), image, true);
    FILE* o = fopen(out, "wb");
    if (!o) throw new Exception(std::string("cannot open ") + out);
    write_vec(o, image);
    fclose(o);
    printf("ilm-render ok: %dx%d, %d spp -> %s\n", loader.Job().Width, loader.Job().Height, loader.Job().Samples,
           loader.Job().Filename.c_str());
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s <plugin.so> host | multi | trace|render <in> <out> <args>\n", argv[0]);
        return 2;
    }
    void* lib = dlopen(argv[1], RTLD_NOW | RTLD_LOCAL);
    if (!lib) {
        fprintf(stderr, "dlopen: %s\n", dlerror());
        return 2;
    }
    auto get = (SYMBOL_SINGLE_GET)dlsym(lib, "GetPlugInSingleton");
    auto dispose = (SYMBOL_SINGLE_DISPOSE)dlsym(lib, "DisposePlugInSingleton");
    if (!get || !dispose) {
        fprintf(stderr, "plug-in symbols missing\n");
        return 2;
    }
    const std::string mode = argv[2];
    if (mode == "multi") {
        int rc = 3;
        try {
            rc = run_multi(lib);
        } catch (Exception* e) {
            fprintf(stderr, "exception: %s\n", e->GetMessage().c_str());
            delete e;
        }
        dlclose(lib);
        return rc;
    }
    IPlugIn* plugin = get();
    int rc = 2;
    try {
        if (mode == "host")
            rc = run_host(plugin);
        else if (mode == "trace" && argc == 6)
            rc = run_trace(plugin, argv[3], argv[4], argv[5]);
        else if (mode == "space" && argc == 6)
            rc = run_space(plugin, argv[3], argv[4]);
        else if (mode == "render" && argc == 6)
            rc = run_render(plugin, argv[3], argv[4], argv[5]);
        else if (mode == "ilm-load" && argc == 5)
            rc = run_ilm_load(argv[3], argv[4]);
        else if (mode == "ilm-render" && argc == 6)
            rc = run_ilm_render(argv[3], argv[4], argv[5]);
        else
            fprintf(stderr, "bad mode or arguments\n");
    } catch (Exception* e) {
        fprintf(stderr, "exception: %s\n", e->GetMessage().c_str());
        delete e;
        rc = 3;
    }
    dispose();
    dlclose(lib);
    return rc;
}
