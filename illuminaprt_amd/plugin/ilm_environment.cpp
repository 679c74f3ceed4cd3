// Scene-script loading for the GPU path (see ilm_environment.hpp).
#include "ilm_environment.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <memory>
#include <set>
#include <sstream>
#include <stack>

namespace IlluminaGpu {

namespace {

bool is_blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }
bool is_alpha(char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || c == '_'; }
bool is_alnum(char c) { return is_alpha(c) || (c >= '0' && c <= '9'); }

std::string trim_ws(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

std::string dir_of(const std::string& path) {
    size_t p = path.find_last_of('/');
    return p == std::string::npos ? std::string() : path.substr(0, p + 1);
}

std::string file_of(const std::string& path) {
    size_t p = path.find_last_of('/');
    return p == std::string::npos ? path : path.substr(p + 1);
}

void normal_of(const Vector3& a, const Vector3& b, const Vector3& c, float n[3]) {
    const float e1[3] = {b.X - a.X, b.Y - a.Y, b.Z - a.Z}, e2[3] = {c.X - a.X, c.Y - a.Y, c.Z - a.Z};
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    float l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (l > 0) {
        n[0] /= l;
        n[1] /= l;
        n[2] /= l;
    }
}

int builder_of(const std::string& container) {
    // WavefrontSceneLoader::LoadGeometry (Scene/WavefrontSceneLoader.cpp:354-368):
    // the Container names a shape factory; anything else gives a KDTreeMesh.
    if (container == "BVHMesh" || container == "BasicMesh") return ILM_BUILDER_SAH;
    if (container == "KDTreeMeshEx") return ILM_BUILDER_KDTREE_EX;
    return ILM_BUILDER_KDTREE;
}

}  // namespace

// ---------------------------------------------------------------- lexer

// Lexer::ExtractToken (System/Lexer.cpp:27-114)
bool Lexer::ReadToken(LexerToken& token) {
    enum State { Start, Name, Value, Comment } state = Start;
    std::string name, value;
    char ch;
    token = LexerToken();
    while (m_in->get(ch)) {
        switch (state) {
        case Start:
            if (is_blank(ch)) {
            } else if (ch == '#') {
                state = Comment;
            } else if (ch == '{') {
                token.Type = LexerToken::LeftCurly;
                token.Value = "{";
                return true;
            } else if (ch == '}') {
                token.Type = LexerToken::RightCurly;
                token.Value = "}";
                return true;
            } else if (ch == '=') {
                state = Value;
                value.clear();
            } else if (is_alpha(ch)) {
                state = Name;
                name = ch;
            }
            break;
        case Name:
            if (is_blank(ch)) {
                token.Type = LexerToken::PropertyName;
                token.Value = name;
                return true;
            }
            if (is_alnum(ch)) name += ch;
            break;
        case Value:
            if (ch == '\r' || ch == '\n') {
                token.Type = LexerToken::PropertyValue;
                token.Value = trim_ws(value);
                return true;
            }
            value += ch;
            break;
        case Comment:
            if (ch == '\n') state = Start;
            break;
        }
    }
    if (state != Start) throw new Exception("Parsing error : Unexpected end of file!");
    token.Type = LexerToken::EndOfStream;
    return false;
}

// ---------------------------------------------------------------- parse tree

bool ParseNode::FindByName(const std::string& name, std::vector<ParseNode*>& out) const {
    out.clear();
    for (ParseNode* c : Children)
        if (c->Name == name) out.push_back(c);
    return !out.empty();
}

bool ParseNode::GetArgumentMap(ArgumentMap& args) const {
    std::map<std::string, std::string> m;
    for (ParseNode* c : Children)
        if (c->Type == Leaf) m[c->Name] = c->Value;
    args = ArgumentMap(m);
    return !m.empty();
}

EnvironmentLoader::~EnvironmentLoader() {
    for (ParseNode* n : m_nodes) delete n;
}

// EnvironmentLoader::Parse (Scene/EnvironmentLoader.cpp:163-213): a name
// followed by '{' opens an internal node, a name followed by any other token
// is a leaf holding that token's text, '}' closes the current node. A name
// at the end of the stream ends parsing with no sections read.
void EnvironmentLoader::Parse(std::istream& in) {
    for (ParseNode* n : m_nodes) delete n;
    m_nodes.clear();
    m_root = ParseNode();
    Lexer lexer(&in);
    std::stack<ParseNode*> stack;
    stack.push(&m_root);
    LexerToken name, value;
    while (lexer.ReadToken(name)) {
        if (name.Type == LexerToken::PropertyName) {
            if (!lexer.ReadToken(value)) {
                m_root.Children.clear();
                Warn("a name at the end of the script: no sections were read");
                return;
            }
            ParseNode* node = new ParseNode();
            m_nodes.push_back(node);
            node->Name = name.Value;
            if (value.Type == LexerToken::LeftCurly) {
                node->Type = ParseNode::Internal;
                stack.top()->Children.push_back(node);
                stack.push(node);
            } else {
                node->Type = ParseNode::Leaf;
                node->Value = value.Value;
                stack.top()->Children.push_back(node);
            }
        } else if (name.Type == LexerToken::RightCurly) {
            if (stack.size() == 1) throw new Exception("Parsing error : unbalanced '}'");
            stack.pop();
        }
    }
}

bool EnvironmentLoader::NodeList(const std::string& category, const std::string& instance,
                                 std::vector<ParseNode*>& out) const {
    // EnvironmentLoader::GetNodeList (:217-244)
    std::vector<ParseNode*> cats, inst;
    out.clear();
    if (!m_root.FindByName(category, cats)) return false;
    for (ParseNode* c : cats)
        if (c->FindByName(instance, inst)) out.insert(out.end(), inst.begin(), inst.end());
    return !out.empty();
}

std::string EnvironmentLoader::Resolve(const std::string& file) const {
    std::ifstream f(file);
    if (f || file.empty() || file[0] == '/' || m_base.empty()) return file;
    return m_base + file;
}

// Each section: entries without Id or Type are skipped with a warning; a
// failed or duplicate instance is reported and skipped (the reference
// catches the factory's exception and prints it, :249-705).
template <class T>
void EnvironmentLoader::Each(const std::string& category, const std::string& instance, const char* label,
                             std::map<std::string, T>& into,
                             T (EnvironmentLoader::*make)(const std::string&, const ArgumentMap&)) {
    std::vector<ParseNode*> nodes;
    if (!NodeList(category, instance, nodes)) return;
    for (ParseNode* n : nodes) {
        ArgumentMap args;
        n->GetArgumentMap(args);
        std::string id, type;
        if (!(args.ContainsArgument("Id") && args.ContainsArgument("Type"))) {
            Warn(std::string("[") + label + "] entry without Id or Type ignored");
            continue;
        }
        args.GetArgument("Id", id);
        args.GetArgument("Type", type);
        if (into.count(id)) {
            Warn(std::string("[") + label + "] " + id + ": Instance name must be unique!");
            continue;
        }
        try {
            into[id] = (this->*make)(type, args);
        } catch (Exception* e) {
            Warn(std::string("[") + label + "] " + id + ": " + e->GetMessage());
            delete e;
        }
    }
}

// ---------------------------------------------------------------- sections

// ShapeFactories.h:140-190 (Quad), :74-115 (Triangle), EnvironmentLoader.cpp:593-640
// (WavefrontModel and the mesh containers).
EnvironmentLoader::Shape EnvironmentLoader::MakeShape(const std::string& type, const ArgumentMap& args) {
    Shape s;
    auto add_vertex = [&](const Vector3& p, const float n[3]) {
        s.pos.insert(s.pos.end(), {p.X, p.Y, p.Z});
        s.nrm.insert(s.nrm.end(), {n[0], n[1], n[2]});
    };
    if (type == "WavefrontModel") {
        std::string file, container, id, group;
        args.GetArgument("Filename", file);
        args.GetArgument("Container", container);
        args.GetArgument("Id", id);
        file = Resolve(trim_ws(file));
        ilm_obj* o = nullptr;
        Check(ilm_obj_load(file.c_str(), &o));
        std::unique_ptr<ilm_obj, int (*)(ilm_obj*)> guard(o, ilm_obj_destroy);
        int64_t nv = 0, nt = 0;
        int32_t nmat = 0, nl = 0;
        Check(ilm_obj_counts(o, &nv, &nt, &nmat, &nl));
        s.pos.resize(nv * 3);
        s.nrm.resize(nv * 3);
        s.idx.resize(nt * 3);
        s.group.resize(nt);
        std::vector<float> mats(nmat * 8);
        Check(ilm_obj_arrays(o, s.pos.data(), s.nrm.data(), s.idx.data(), s.group.data(), nullptr, mats.data(),
                             nullptr));
        std::vector<int32_t> illum(nmat);
        char lib[512];
        Check(ilm_obj_material_info(o, illum.data(), lib, sizeof lib));
        s.builder = builder_of(container);
        // The material group (WavefrontSceneLoader.cpp:137-146, :304-330): Matte
        // from Kd, glass (illum 4) and mirror (illum 5) are not drawn by the
        // GPU path. Emission (Ke) is not read, as in the reference.
        Material g;
        g.group = true;
        for (int k = 0; k < nmat; ++k) {
            const float rec[8] = {mats[8 * k], mats[8 * k + 1], mats[8 * k + 2], 0, 0, 0, 0, 0};
            g.records.insert(g.records.end(), rec, rec + 8);
            g.member_unsupported.push_back(illum[k] == 4   ? "glass materials are not drawn by the GPU path"
                                           : illum[k] == 5 ? "mirror materials are not drawn by the GPU path"
                                                           : "");
        }
        group = lib[0] ? std::string(lib) : file_of(file) + ".materials";
        args.GetArgument("MaterialGroupId", group);
        if (m_materials.count(group))
            Warn("[Shape] " + id + ": material group " + group + " already exists");
        else
            m_materials[group] = g;
        return s;
    }
    if (type == "Quad") {
        std::vector<Vector3> v;
        args.GetArgument("Vertices", v);
        if (v.size() != 4) throw new Exception("QuadMeshShapeFactory : Incorrect number of vertices.");
        // Quad::Quad (Shape/Quad.cpp:11-29): triangles (v1, v3, v2) and (v1, v4, v3)
        float n[3];
        normal_of(v[0], v[2], v[1], n);
        for (const Vector3& p : v) add_vertex(p, n);
        s.idx = {0, 2, 1, 0, 3, 2};
        s.group = {0, 0};
        return s;
    }
    if (type == "Triangle") {
        Vector3 p[3];
        if (!(args.GetArgument("P0", p[0]) && args.GetArgument("P1", p[1]) && args.GetArgument("P2", p[2])))
            throw new Exception("Invalid arguments to TriangleShapeFactory!");
        float n[3];
        normal_of(p[0], p[1], p[2], n);
        for (const Vector3& q : p) add_vertex(q, n);
        s.idx = {0, 1, 2};
        s.group = {0};
        return s;
    }
    if (type == "BVHMesh" || type == "KDTreeMesh" || type == "KDTreeMeshEx" || type == "BasicMesh") {
        s.builder = builder_of(type);  // an empty mesh container
        return s;
    }
    if (type == "Sphere") {
        s.unsupported = "spheres are not triangle meshes; the GPU path traces triangles";
        return s;
    }
    if (type == "PersistentMesh") {
        // ShapeFactories.h:337-370: TrunkName names <trunk>.ver / .tri / .tre.
        // The geometry and group ids are read; the script's space is
        // flattened and gets one tree, built as a KDTreeMesh.
        std::string trunk;
        if (!args.GetArgument("TrunkName", trunk)) throw new Exception("PersistentMesh needs TrunkName");
        trunk = Resolve(trim_ws(trunk) + ".ver");
        trunk = trunk.substr(0, trunk.size() - 4);
        ilm_persistent* p = nullptr;
        Check(ilm_persistent_load(trunk.c_str(), &p));
        std::unique_ptr<ilm_persistent, int (*)(ilm_persistent*)> guard(p, ilm_persistent_destroy);
        int64_t nv = 0, nt = 0;
        Check(ilm_persistent_counts(p, &nv, &nt, nullptr, nullptr));
        s.pos.resize(3 * nv);
        s.nrm.resize(3 * nv);
        s.idx.resize(3 * nt);
        s.group.resize(nt);
        Check(ilm_persistent_arrays(p, s.pos.data(), s.nrm.data(), s.idx.data(), s.group.data()));
        s.builder = ILM_BUILDER_KDTREE;
        return s;
    }
    throw new Exception("No factory registered under specified name!");
}

// MaterialFactories.h:49-75 (Matte); Mirror, Glass and Group are recorded as not drawable.
EnvironmentLoader::Material EnvironmentLoader::MakeMaterial(const std::string& type, const ArgumentMap& args) {
    Material m;
    if (type == "Matte") {
        Vector3 r;
        if (!args.GetArgument("Reflectivity", r)) throw new Exception("Invalid arguments to MatteMaterialFactory!");
        m.records = {r.X, r.Y, r.Z, 0, 0, 0, 0, 0};
        m.member_unsupported.push_back("");
        return m;
    }
    if (type == "Mirror" || type == "Glass") {
        m.unsupported = type + " materials are not drawn by the GPU path";
        return m;
    }
    if (type == "Group") {
        m.unsupported = "material groups come from WavefrontModel shapes";
        return m;
    }
    throw new Exception("No factory registered under specified name!");
}

// LightFactories.h:193-229 (DiffuseArea: Power is the emitted radiance) and
// the Shape binding of EnvironmentLoader::ParseLights (:336-341).
EnvironmentLoader::Light EnvironmentLoader::MakeLight(const std::string& type, const ArgumentMap& args) {
    Light l;
    if (type == "DiffuseArea") {
        Vector3 p;
        if (args.GetArgument("Power", p)) {
            l.power[0] = p.X;
            l.power[1] = p.Y;
            l.power[2] = p.Z;
        }
        std::string shape;
        if (args.GetArgument("Shape", shape)) {
            if (!m_shapes.count(shape)) Warn("[Light] Cannot assign shape instance to area light: " + shape);
            else l.shape = shape;
        }
        return l;
    }
    if (type == "Point" || type == "Spot" || type == "Directional" || type == "InfiniteArea") {
        l.unsupported = type + " lights are not drawn by the GPU path (area lights only)";
        return l;
    }
    throw new Exception("No factory registered under specified name!");
}

// CameraFactories.h:40-78. The frame at Distance d with half-height
// tan(fov/2) spans the same directions as the frame at distance 1 with
// half-height tan(fov/2) / d, so d folds into the field of view.
EnvironmentLoader::Camera EnvironmentLoader::MakeCamera(const std::string& type, const ArgumentMap& args) {
    Camera c;
    if (type == "Perspective") {
        float aspect = 1.0f, fov = 60.0f, distance = 1.0f;
        Vector3 pos(0, 0, -1), target(0, 0, 0), up(0, 1, 0);
        args.GetArgument("Aspect", aspect);
        args.GetArgument("Fov", fov);
        args.GetArgument("Position", pos);
        args.GetArgument("Distance", distance);
        args.GetArgument("Target", target);
        args.GetArgument("Up", up);
        if (!(distance > 0)) throw new Exception("Perspective camera: Distance must be positive");
        const float* src[3] = {&pos.X, &target.X, &up.X};
        float* dst[3] = {c.cam.pos, c.cam.target, c.cam.up};
        for (int k = 0; k < 3; ++k) std::memcpy(dst[k], src[k], 3 * sizeof(float));
        c.cam.fov_deg = distance == 1.0f
                            ? fov
                            : (float)(2.0 * std::atan(std::tan(fov * 3.14159265358979323846 / 360.0) / distance) *
                                      180.0 / 3.14159265358979323846);
        c.cam.aspect = aspect;
        return c;
    }
    if (type == "ThinLens") {
        c.unsupported = "the thin-lens camera is not drawn by the GPU path";
        return c;
    }
    throw new Exception("No factory registered under specified name!");
}

// IntegratorFactories.h:30-68 (PathTracing: RayDepth 6, ShadowRays 1, Epsilon 1e-4).
EnvironmentLoader::Integrator EnvironmentLoader::MakeIntegrator(const std::string& type, const ArgumentMap& args) {
    Integrator it;
    if (type == "PathTracing") {
        int shadow = 1;
        args.GetArgument("RayDepth", it.depth);
        args.GetArgument("ShadowRays", shadow);
        args.GetArgument("Epsilon", it.eps);
        if (shadow != 1) it.unsupported = "PathTracing with ShadowRays other than 1";
        if (it.depth < 1) it.unsupported = "PathTracing with RayDepth below 1";
        return it;
    }
    if (type == "IC" || type == "MLIC" || type == "IGI" || type == "Whitted") {
        it.unsupported = "the " + type + " integrator is not on the GPU path (PathTracing only)";
        return it;
    }
    throw new Exception("No factory registered under specified name!");
}

// RendererFactories.h:23-56 (Basic: Samples 1) and the bindings of
// EnvironmentLoader::ParseRenderers (:667-704).
EnvironmentLoader::Renderer EnvironmentLoader::MakeRenderer(const std::string& type, const ArgumentMap& args) {
    Renderer r;
    args.GetArgument("Integrator", r.integrator);
    args.GetArgument("Device", r.device);
    args.GetArgument("Filter", r.filter);
    if (type == "Basic") {
        args.GetArgument("Samples", r.samples);
        if (r.samples < 1) r.unsupported = "Samples below 1";
        return r;
    }
    if (type == "TimeConstrained" || type == "Disparity" || type == "Distributed") {
        r.unsupported = "the " + type + " renderer is not on the GPU path (Basic only)";
        return r;
    }
    throw new Exception("No factory registered under specified name!");
}

// DeviceFactories.h:330-378 (Image: 640 x 480, "result.ppm"; the reference
// writes PPM whatever Format says; Format = PFM selects PFM here).
EnvironmentLoader::Device EnvironmentLoader::MakeDevice(const std::string& type, const ArgumentMap& args) {
    Device d;
    args.GetArgument("Width", d.width);
    args.GetArgument("Height", d.height);
    args.GetArgument("Format", d.format);
    args.GetArgument("Filename", d.filename);
    if (type == "Image" || type == "BufferedImage") {
        if (d.width < 1 || d.height < 1) d.unsupported = "image size not positive";
        return d;
    }
    if (type == "SharedMemory" || type == "GLDisplay" || type == "Display" || type == "Video" || type == "RTP") {
        d.unsupported = "the " + type + " device is not written by the GPU path (Image only)";
        return d;
    }
    throw new Exception("No factory registered under specified name!");
}

// Samplers, filters and textures: names only. The GPU path draws its samples
// from its own counter-based streams and uses the box filter.
std::string EnvironmentLoader::MakeName(const std::string& type, const ArgumentMap&) {
    static const std::set<std::string> samplers = {"Random", "Jitter", "Multijitter", "LowDiscrepancy",
                                                   "PrecomputedRandom", "PrecomputedHalton", "PrecomputedSobol"};
    if (samplers.count(type) || type == "Box") return "";
    if (type == "Tent") return "the tent filter is not applied by the GPU path (box only)";
    if (type == "MappedFile" || type == "Image" || type == "Noise" || type == "Marble")
        return "textures are not read by the GPU path";
    throw new Exception("No factory registered under specified name!");
}

// ---------------------------------------------------------------- environment

// EnvironmentLoader::ParseEnvironment (:711-912) resolved into one flattened
// GPU scene: every primitive of the Basic space, with its WorldTransform,
// material (or the material group of its mesh) and, for Emissive
// primitives, its light.
void EnvironmentLoader::ParseEnvironment() {
    auto need = [](bool ok, const std::string& what) {
        if (!ok) throw new Exception("EnvironmentLoader :: " + what);
    };
    auto usable = [](const std::string& unsupported, const std::string& what) {
        if (!unsupported.empty()) throw new Exception(what + ": " + unsupported);
    };
    std::vector<ParseNode*> env;
    need(m_root.FindByName("Environment", env), "Missing Environment block!");
    ArgumentMap ea;
    env[0]->GetArgumentMap(ea);
    std::string rid;
    need(ea.GetArgument("Renderer", rid), "Missing Renderer within Environment block!");
    need(m_renderers.count(rid) != 0, "Instance does not exist! (renderer " + rid + ")");
    const Renderer& r = m_renderers[rid];
    usable(r.unsupported, "renderer " + rid);
    need(m_integrators.count(r.integrator) != 0, "Unable to bind integrator [" + r.integrator + "]");
    const Integrator& in = m_integrators[r.integrator];
    usable(in.unsupported, "integrator " + r.integrator);
    need(m_devices.count(r.device) != 0, "Unable to bind device [" + r.device + "]");
    const Device& dev = m_devices[r.device];
    usable(dev.unsupported, "device " + r.device);
    if (m_filters.count(r.filter))
        usable(m_filters[r.filter], "filter " + r.filter);
    else
        Warn("[Renderer] Unable to bind filter [" + r.filter + "]; the box filter is used");

    std::vector<ParseNode*> scene;
    need(env[0]->FindByName("Scene", scene), "Missing Scene block within Environment block!");
    ArgumentMap sa;
    scene[0]->GetArgumentMap(sa);
    std::string sid, cid;
    need(sa.GetArgument("Sampler", sid), "Missing Sampler within Scene block!");
    need(m_samplers.count(sid) != 0, "Instance does not exist! (sampler " + sid + ")");
    need(sa.GetArgument("Camera", cid), "Missing Camera within Scene block!");
    need(m_cameras.count(cid) != 0, "Instance does not exist! (camera " + cid + ")");
    usable(m_cameras[cid].unsupported, "camera " + cid);
    std::vector<std::string> light_ids;
    sa.GetArgument("Lights", light_ids);
    for (const std::string& l : light_ids) {
        need(m_lights.count(l) != 0, "Instance does not exist! (light " + l + ")");
        usable(m_lights[l].unsupported, "light " + l);
    }

    std::vector<ParseNode*> space, prims, nodes;
    need(scene[0]->FindByName("Space", space), "Missing Space block within Scene block!");
    ArgumentMap spa;
    space[0]->GetArgumentMap(spa);
    std::string space_id, space_type;
    need(spa.GetArgument("Id", space_id) && spa.GetArgument("Type", space_type), "Space needs Id and Type");
    need(space_type == "Basic", "No factory registered under specified name! (space type " + space_type + ")");
    need(space[0]->FindByName("Primitives", prims) && prims[0]->FindByName("Primitive", nodes),
         "Missing Primitives within Space block!");

    RenderJob job;
    job.Camera = m_cameras[cid].cam;
    job.Width = dev.width;
    job.Height = dev.height;
    job.Samples = r.samples;
    job.RayDepth = in.depth;
    job.Epsilon = in.eps;
    job.Filename = dev.filename;
    job.Format = dev.format == "PFM" ? ILM_IMAGE_PFM : ILM_IMAGE_PPM;
    job.LightIds = light_ids;
    for (const std::string& l : light_ids) {
        const Light& L = m_lights[l];
        job.Lights.insert(job.Lights.end(), {L.power[0], L.power[1], L.power[2], 0.0f});
    }

    ilm_space* sp = nullptr;
    Check(ilm_space_create(nullptr, &sp));
    std::unique_ptr<ilm_space, int (*)(ilm_space*)> guard(sp, ilm_space_destroy);
    std::map<std::string, int32_t> mat_base;   // material id -> first record
    std::map<std::string, std::string> light_prim;
    std::vector<int32_t> tri_mat;
    bool have_builder = false;
    for (ParseNode* pn : nodes) {
        ArgumentMap pa;
        pn->GetArgumentMap(pa);
        if (!(pa.ContainsArgument("Id") && pa.ContainsArgument("Type"))) {
            Warn("[Primitive] entry without Id or Type ignored");
            continue;
        }
        std::string id, type, mid, shid, lid;
        pa.GetArgument("Id", id);
        pa.GetArgument("Type", type);
        if (type != "Emissive" && type != "Geometry") continue;
        pa.GetArgument("Material", mid);
        pa.GetArgument("Shape", shid);
        need(m_materials.count(mid) != 0, "Instance does not exist! (material " + mid + ")");
        need(m_shapes.count(shid) != 0, "Instance does not exist! (shape " + shid + ")");
        const Material& M = m_materials[mid];
        const Shape& S = m_shapes[shid];
        usable(M.unsupported, "material " + mid);
        usable(S.unsupported, "shape " + shid);
        int32_t light = -1;
        if (type == "Emissive") {
            pa.GetArgument("Light", lid);
            need(m_lights.count(lid) != 0, "Instance does not exist! (light " + lid + ")");
            usable(m_lights[lid].unsupported, "light " + lid);
            auto pos = std::find(light_ids.begin(), light_ids.end(), lid);
            if (pos == light_ids.end())
                throw new Exception("light " + lid + " of primitive " + id +
                                    " is not in the Scene's Lights list (the GPU path samples every emitter)");
            if (light_prim.count(lid))
                throw new Exception("light " + lid + " is bound to primitives " + light_prim[lid] + " and " + id);
            if (!m_lights[lid].shape.empty() && m_lights[lid].shape != shid)
                throw new Exception("light " + lid + " samples shape " + m_lights[lid].shape +
                                    " but its primitive " + id + " has shape " + shid);
            light_prim[lid] = id;
            light = (int32_t)(pos - light_ids.begin());
        }
        Primitive t;  // WorldTransform: identity unless given (:842-858)
        float rot[9];
        Vector3 v;
        if (pa.GetArgument("Rotation", rot)) std::memcpy(t.Rotation, rot, sizeof rot);
        if (pa.GetArgument("Translation", v)) {
            t.Translation[0] = v.X;
            t.Translation[1] = v.Y;
            t.Translation[2] = v.Z;
        }
        if (pa.GetArgument("Scaling", v)) {
            t.Scaling[0] = v.X;
            t.Scaling[1] = v.Y;
            t.Scaling[2] = v.Z;
        }
        if (!mat_base.count(mid)) {
            mat_base[mid] = (int32_t)(job.Materials.size() / 8);
            job.Materials.insert(job.Materials.end(), M.records.begin(), M.records.end());
        }
        const int64_t nt = (int64_t)S.idx.size() / 3;
        for (int64_t k = 0; k < nt; ++k) {
            int32_t g = M.group ? S.group[k] : 0;
            if (g < 0 || g >= (int32_t)M.member_unsupported.size()) g = 0;
            usable(M.member_unsupported[g], "material " + mid + " member " + std::to_string(g));
            tri_mat.push_back(mat_base[mid] + g);
        }
        if (!have_builder && nt > 0) {
            job.Builder = S.builder;
            have_builder = true;
        }
        int32_t prim = 0;
        if (nt > 0)
            Check(ilm_space_add(sp, S.pos.data(), (int64_t)S.pos.size() / 3, S.idx.data(), nt, S.nrm.data(),
                                t.Rotation, t.Scaling, t.Translation, 0, light, &prim));
        job.PrimitiveIds.push_back(id);
        job.PrimitiveFirstTriangle.push_back((int64_t)tri_mat.size() - nt);
    }
    job.PrimitiveFirstTriangle.push_back((int64_t)tri_mat.size());
    for (const std::string& l : light_ids)
        if (!light_prim.count(l))
            throw new Exception("light " + l + " is not bound to an Emissive primitive (the GPU path samples "
                                "emitters through their triangles)");
    need(!tri_mat.empty(), "the space holds no triangles");
    ilm_bvh_params bp;
    std::memset(&bp, 0, sizeof bp);
    bp.builder = ILM_BUILDER_SAH;
    bp.width = 4;
    Check(ilm_space_build(sp, &bp));
    int64_t np = 0, nv = 0, nt = 0;
    Check(ilm_space_size(sp, &np, &nv, &nt));
    job.Positions.resize(nv * 3);
    job.Normals.resize(nv * 3);
    job.Indices.resize(nt * 3);
    job.TriLight.resize(nt);
    Check(ilm_space_geometry(sp, job.Positions.data(), job.Indices.data()));
    Check(ilm_space_shading(sp, job.Normals.data(), nullptr, job.TriLight.data()));
    job.TriMaterial = tri_mat;
    m_job = job;
}

void EnvironmentLoader::Load(std::istream& in, const std::string& base_dir) {
    m_base = base_dir;
    m_warnings.clear();
    m_shapes.clear();
    m_materials.clear();
    m_lights.clear();
    m_cameras.clear();
    m_integrators.clear();
    m_renderers.clear();
    m_devices.clear();
    m_filters.clear();
    m_samplers.clear();
    m_textures.clear();
    Parse(in);
    // the order of EnvironmentLoader::Parse (:198-208)
    Each("Textures", "Texture", "Texture", m_textures, &EnvironmentLoader::MakeName);
    Each("Shapes", "Shape", "Shape", m_shapes, &EnvironmentLoader::MakeShape);
    Each("Materials", "Material", "Material", m_materials, &EnvironmentLoader::MakeMaterial);
    Each("Lights", "Light", "Light", m_lights, &EnvironmentLoader::MakeLight);
    Each("Cameras", "Camera", "Camera", m_cameras, &EnvironmentLoader::MakeCamera);
    Each("Filters", "Filter", "Filter", m_filters, &EnvironmentLoader::MakeName);
    Each("Samplers", "Sampler", "Sampler", m_samplers, &EnvironmentLoader::MakeName);
    Each("Devices", "Device", "Device", m_devices, &EnvironmentLoader::MakeDevice);
    Each("Integrators", "Integrator", "Integrator", m_integrators, &EnvironmentLoader::MakeIntegrator);
    Each("Renderers", "Renderer", "Renderer", m_renderers, &EnvironmentLoader::MakeRenderer);
    ParseEnvironment();
}

bool EnvironmentLoader::Load(const std::string& filename) {
    std::ifstream f(filename);
    if (!f.is_open()) return false;
    Load(f, dir_of(filename));
    return true;
}

// ---------------------------------------------------------------- rendering

void RenderJob::Render(uint32_t seed, std::vector<float>& image, bool write) const {
    ilm_ctx* ctx = GpuContext();
    GpuBVHMesh mesh(ctx, 10, 35, Builder);
    const int64_t nv = (int64_t)Positions.size() / 3, nt = (int64_t)Indices.size() / 3;
    for (int64_t v = 0; v < nv; ++v) mesh.AddVertex(Vector3(Positions[3 * v], Positions[3 * v + 1], Positions[3 * v + 2]));
    for (int64_t t = 0; t < nt; ++t) mesh.AddIndexedTriangle(Indices[3 * t], Indices[3 * t + 1], Indices[3 * t + 2]);
    mesh.Compile();
    GpuPathTracing pt(ctx, RayDepth, Epsilon);
    pt.Initialise(&mesh, Normals, TriMaterial, TriLight, Materials, Lights);
    CameraArgs cam;
    cam.Position = Vector3(Camera.pos[0], Camera.pos[1], Camera.pos[2]);
    cam.Target = Vector3(Camera.target[0], Camera.target[1], Camera.target[2]);
    cam.Up = Vector3(Camera.up[0], Camera.up[1], Camera.up[2]);
    cam.Fov = Camera.fov_deg;
    cam.Aspect = Camera.aspect;
    std::vector<float> frame;
    pt.RenderRegion(cam, Width, Height, Samples, seed, frame);
    // ImageDevice::Set writes frame pixel (x, y) at (w-1-x, h-1-y) (Device/ImageDevice.cpp:137-148)
    image.assign(frame.size(), 0.0f);
    const size_t W = (size_t)Width, H = (size_t)Height;
    for (size_t y = 0; y < H; ++y)
        for (size_t x = 0; x < W; ++x)
            for (int c = 0; c < 3; ++c)
                image[3 * ((H - 1 - y) * W + (W - 1 - x)) + c] = frame[3 * (y * W + x) + c];
    if (write) Check(ilm_image_write(Filename.c_str(), image.data(), Width, Height, Format));
}

}  // namespace IlluminaGpu
