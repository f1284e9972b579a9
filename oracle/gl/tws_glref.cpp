// tws_glref.cpp -- runs the reference's own GLSL compute shaders on the CPU
// through Mesa's llvmpipe, to pin the CPU restatement (oracle/tws_oracle.c).
//
// TEST INFRASTRUCTURE ONLY; it never ships and never runs on the GPU side.
// The shaders are read at run time from the reference checkout
// (terrainwatersim/shader/{flowUpdate,flowApply,waterBrush}.comp and their
// includes); nothing of them is copied.
//
// Context: Mesa's swrast DRI driver (swrast_dri.so) driven directly through
// the public DRI interface (GL/internal/dri_interface.h): DRI_SWRast
// createNewScreen2 -> createContextAttribs(OpenGL core 4.3) -> DRI_Core
// bindContext with no drawables (compute only, no framebuffer).  GL entry
// points come from the shared glapi dispatch (_glapi_get_proc_address).
//
// Reference behaviour reproduced here:
//   - shader text: "#include" spliced recursively, each file once
//     (glEasy/gl/ShaderObject.cpp:77-153), no defines for these programs
//     (Terrain.cpp:68-74);
//   - uniform blocks use the "shared" layout, so member offsets are read by
//     reflection as glEasy's UniformBuffer does (UniformBuffer.cpp:59-120);
//   - per step: flowUpdate with TerrainData bound READ at unit 0 and Flow
//     READ_WRITE at unit 1, dispatch (W/16, H/16); then flowApply with
//     TerrainData READ_WRITE, Flow READ, FlowMap WRITE at unit 2, same
//     dispatch (Terrain.cpp:253-265); brush: TerrainData at unit 0, BrushInfo
//     at binding 7, dispatch (W/32, H/32) (Terrain.cpp:150-168).
// Harness-side fixes (the shaders stay untouched): a
// glMemoryBarrier(GL_SHADER_IMAGE_ACCESS_BARRIER_BIT) between dependent
// dispatches (the reference issues none), and the brush binds TerrainData
// READ_WRITE (the reference binds it READ and writes it).

#include <GL/glcorearb.h>
#include <GL/internal/dri_interface.h>
#include <dlfcn.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <set>
#include <sstream>
#include <string>
#include <vector>

namespace {

std::string g_err;
std::string g_renderer;

// ---- DRI context ------------------------------------------------------------
const __DRIcoreExtension* g_core = nullptr;
const __DRIswrastExtension* g_swrast = nullptr;
__DRIscreen* g_screen = nullptr;
__DRIcontext* g_ctx = nullptr;
void* g_glapi = nullptr;
void* g_driver = nullptr;
typedef void* (*GetProcFn)(const char*);
GetProcFn g_getproc = nullptr;

void drawable_info(__DRIdrawable*, int* x, int* y, int* w, int* h, void*) {
    *x = *y = 0;
    *w = *h = 1;
}
void put_image(__DRIdrawable*, int, int, int, int, int, char*, void*) {}
void get_image(__DRIdrawable*, int, int, int, int, char* data, void*) {
    if (data) memset(data, 0, 4);
}

__DRIswrastLoaderExtension g_loader_ext = {
    {__DRI_SWRAST_LOADER, 1}, drawable_info, put_image, get_image,
    nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
const __DRIextension* g_loader_exts[] = {&g_loader_ext.base, nullptr};

// ---- GL entry points ----------------------------------------------------------
#define GLFN(type, name) type name = nullptr;
#define GL_FUNCS(X)                                                        \
    X(PFNGLGETERRORPROC, glGetError)                                       \
    X(PFNGLGETSTRINGPROC, glGetString)                                     \
    X(PFNGLCREATESHADERPROC, glCreateShader)                               \
    X(PFNGLSHADERSOURCEPROC, glShaderSource)                               \
    X(PFNGLCOMPILESHADERPROC, glCompileShader)                             \
    X(PFNGLGETSHADERIVPROC, glGetShaderiv)                                 \
    X(PFNGLGETSHADERINFOLOGPROC, glGetShaderInfoLog)                       \
    X(PFNGLCREATEPROGRAMPROC, glCreateProgram)                             \
    X(PFNGLATTACHSHADERPROC, glAttachShader)                               \
    X(PFNGLLINKPROGRAMPROC, glLinkProgram)                                 \
    X(PFNGLGETPROGRAMIVPROC, glGetProgramiv)                               \
    X(PFNGLGETPROGRAMINFOLOGPROC, glGetProgramInfoLog)                     \
    X(PFNGLUSEPROGRAMPROC, glUseProgram)                                   \
    X(PFNGLGENTEXTURESPROC, glGenTextures)                                 \
    X(PFNGLDELETETEXTURESPROC, glDeleteTextures)                           \
    X(PFNGLBINDTEXTUREPROC, glBindTexture)                                 \
    X(PFNGLTEXSTORAGE2DPROC, glTexStorage2D)                               \
    X(PFNGLTEXSUBIMAGE2DPROC, glTexSubImage2D)                             \
    X(PFNGLGETTEXIMAGEPROC, glGetTexImage)                                 \
    X(PFNGLPIXELSTOREIPROC, glPixelStorei)                                 \
    X(PFNGLBINDIMAGETEXTUREPROC, glBindImageTexture)                       \
    X(PFNGLGENBUFFERSPROC, glGenBuffers)                                   \
    X(PFNGLDELETEBUFFERSPROC, glDeleteBuffers)                             \
    X(PFNGLBINDBUFFERPROC, glBindBuffer)                                   \
    X(PFNGLBUFFERDATAPROC, glBufferData)                                   \
    X(PFNGLBINDBUFFERBASEPROC, glBindBufferBase)                           \
    X(PFNGLGETUNIFORMBLOCKINDEXPROC, glGetUniformBlockIndex)               \
    X(PFNGLGETACTIVEUNIFORMBLOCKIVPROC, glGetActiveUniformBlockiv)         \
    X(PFNGLGETUNIFORMINDICESPROC, glGetUniformIndices)                     \
    X(PFNGLGETACTIVEUNIFORMSIVPROC, glGetActiveUniformsiv)                 \
    X(PFNGLDISPATCHCOMPUTEPROC, glDispatchCompute)                         \
    X(PFNGLMEMORYBARRIERPROC, glMemoryBarrier)                             \
    X(PFNGLGENERATEMIPMAPPROC, glGenerateMipmap)                           \
    X(PFNGLFINISHPROC, glFinish)
GL_FUNCS(GLFN)

bool load_gl() {
#define LOADFN(type, name)                                            \
    name = reinterpret_cast<type>(g_getproc(#name));                  \
    if (!name) { g_err = "missing GL entry point " #name; return false; }
    GL_FUNCS(LOADFN)
#undef LOADFN
    return true;
}

bool gl_ok(const char* what) {
    GLenum e = glGetError();
    if (e != GL_NO_ERROR) {
        char buf[128];
        snprintf(buf, sizeof(buf), "%s: GL error 0x%x", what, e);
        g_err = buf;
        return false;
    }
    return true;
}

// ---- shader loading (include expansion as ShaderObject.cpp:77-153) -------------
bool read_file(const std::string& path, std::string* out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    *out = ss.str();
    return true;
}

bool expand(const std::string& dir, const std::string& name, std::set<std::string>& seen,
            std::string* out) {
    std::string src;
    const std::string path = dir + "/" + name;
    if (!read_file(path, &src)) {
        g_err = "cannot read shader " + path;
        return false;
    }
    seen.insert(name);
    size_t pos;
    while ((pos = src.find("#include")) != std::string::npos) {
        const size_t q1 = src.find('"', pos);
        const size_t q2 = q1 == std::string::npos ? q1 : src.find('"', q1 + 1);
        if (q2 == std::string::npos) {
            g_err = "bad #include in " + path;
            return false;
        }
        const std::string inc = src.substr(q1 + 1, q2 - q1 - 1);
        std::string body;
        if (!seen.count(inc) && !expand(dir, inc, seen, &body)) return false;
        src.replace(pos, q2 + 1 - pos, body);
    }
    *out = src;
    return true;
}

GLuint build_program(const std::string& dir, const std::string& name) {
    std::set<std::string> seen;
    std::string src;
    if (!expand(dir, name, seen, &src)) return 0;
    GLuint sh = glCreateShader(GL_COMPUTE_SHADER);
    const char* p = src.c_str();
    glShaderSource(sh, 1, &p, nullptr);
    glCompileShader(sh);
    GLint ok = 0;
    glGetShaderiv(sh, GL_COMPILE_STATUS, &ok);
    if (!ok) {
        char log[4096];
        glGetShaderInfoLog(sh, sizeof(log), nullptr, log);
        g_err = "compile " + name + ": " + log;
        return 0;
    }
    GLuint prog = glCreateProgram();
    glAttachShader(prog, sh);
    glLinkProgram(prog);
    glGetProgramiv(prog, GL_LINK_STATUS, &ok);
    if (!ok) {
        char log[4096];
        glGetProgramInfoLog(prog, sizeof(log), nullptr, log);
        g_err = "link " + name + ": " + log;
        return 0;
    }
    return prog;
}

// Uniform block member offsets by reflection ("shared" layout).
struct Block {
    GLuint buf = 0;
    GLint size = 0;
    std::vector<GLint> offs;
};

bool reflect_block(GLuint prog, const char* block, const std::vector<const char*>& names,
                   Block* out) {
    GLuint bi = glGetUniformBlockIndex(prog, block);
    if (bi == GL_INVALID_INDEX) {
        g_err = std::string("no uniform block ") + block;
        return false;
    }
    glGetActiveUniformBlockiv(prog, bi, GL_UNIFORM_BLOCK_DATA_SIZE, &out->size);
    std::vector<GLuint> idx(names.size());
    glGetUniformIndices(prog, (GLsizei)names.size(), names.data(), idx.data());
    for (GLuint i : idx)
        if (i == GL_INVALID_INDEX) {
            g_err = std::string("missing member in ") + block;
            return false;
        }
    out->offs.resize(names.size());
    glGetActiveUniformsiv(prog, (GLsizei)idx.size(), idx.data(), GL_UNIFORM_OFFSET,
                          out->offs.data());
    glGenBuffers(1, &out->buf);
    return gl_ok("reflect_block");
}

void upload_block(const Block& b, const std::vector<float>& vals, GLuint binding) {
    std::vector<unsigned char> data((size_t)b.size, 0);
    for (size_t i = 0; i < vals.size(); ++i) memcpy(&data[(size_t)b.offs[i]], &vals[i], 4);
    glBindBuffer(GL_UNIFORM_BUFFER, b.buf);
    glBufferData(GL_UNIFORM_BUFFER, b.size, data.data(), GL_DYNAMIC_DRAW);
    glBindBufferBase(GL_UNIFORM_BUFFER, binding, b.buf);
}

GLuint g_update = 0, g_apply = 0, g_brush = 0;
Block g_sim_block, g_brush_block;

GLuint make_tex(GLenum fmt, int W, int H) {
    GLuint t = 0;
    glGenTextures(1, &t);
    glBindTexture(GL_TEXTURE_2D, t);
    glTexStorage2D(GL_TEXTURE_2D, 1, fmt, W, H);
    return t;
}

}  // namespace

extern "C" {

const char* glref_last_error(void) { return g_err.c_str(); }
const char* glref_renderer(void) { return g_renderer.c_str(); }

int glref_init(const char* shader_dir, const char* driver_path) {
    if (g_ctx) return 0;
    g_glapi = dlopen("libglapi.so.0", RTLD_NOW | RTLD_GLOBAL);
    if (!g_glapi) { g_err = std::string("dlopen libglapi: ") + dlerror(); return -1; }
    g_getproc = reinterpret_cast<GetProcFn>(dlsym(g_glapi, "_glapi_get_proc_address"));
    if (!g_getproc) { g_err = "no _glapi_get_proc_address"; return -1; }
    g_driver = dlopen(driver_path, RTLD_NOW | RTLD_GLOBAL);
    if (!g_driver) { g_err = std::string("dlopen driver: ") + dlerror(); return -1; }
    typedef const __DRIextension** (*GetExt)(void);
    GetExt get = reinterpret_cast<GetExt>(dlsym(g_driver, "__driDriverGetExtensions_swrast"));
    if (!get) { g_err = "no __driDriverGetExtensions_swrast"; return -1; }
    const __DRIextension** exts = get();
    for (int i = 0; exts[i]; ++i) {
        if (!strcmp(exts[i]->name, __DRI_CORE)) g_core = (const __DRIcoreExtension*)exts[i];
        if (!strcmp(exts[i]->name, __DRI_SWRAST)) g_swrast = (const __DRIswrastExtension*)exts[i];
    }
    if (!g_core || !g_swrast || g_swrast->base.version < 4) {
        g_err = "driver lacks DRI_Core / DRI_SWRast v4";
        return -1;
    }
    const __DRIconfig** configs = nullptr;
    g_screen = g_swrast->createNewScreen2(0, g_loader_exts, exts, &configs, nullptr);
    if (!g_screen) { g_err = "createNewScreen2 failed"; return -1; }
    const uint32_t attribs[] = {__DRI_CTX_ATTRIB_MAJOR_VERSION, 4, __DRI_CTX_ATTRIB_MINOR_VERSION, 3};
    unsigned error = 0;
    g_ctx = g_swrast->createContextAttribs(g_screen, __DRI_API_OPENGL_CORE, nullptr, nullptr, 2,
                                           attribs, &error, nullptr);
    if (!g_ctx && configs && configs[0])
        g_ctx = g_swrast->createContextAttribs(g_screen, __DRI_API_OPENGL_CORE, configs[0],
                                               nullptr, 2, attribs, &error, nullptr);
    if (!g_ctx) {
        g_err = "createContextAttribs failed, error " + std::to_string(error);
        return -1;
    }
    if (!g_core->bindContext(g_ctx, nullptr, nullptr)) { g_err = "bindContext failed"; return -1; }
    if (!load_gl()) return -1;
    const GLubyte* r = glGetString(GL_RENDERER);
    const GLubyte* v = glGetString(GL_VERSION);
    g_renderer = std::string(r ? (const char*)r : "?") + " | " + (v ? (const char*)v : "?");
    const std::string dir(shader_dir);
    g_update = build_program(dir, "flowUpdate.comp");
    if (!g_update) return -1;
    g_apply = build_program(dir, "flowApply.comp");
    if (!g_apply) return -1;
    g_brush = build_program(dir, "waterBrush.comp");
    if (!g_brush) return -1;
    if (!reflect_block(g_update, "SimulationParameters",
                       {"FlowFriction_perStep", "WaterAcceleration_perStep",
                        "CellAreaInv_timeScaled"}, &g_sim_block))
        return -1;
    if (!reflect_block(g_brush, "BrushInfo",
                       {"BrushPositionTexelCor", "BrushIntensity", "BrushSizeSq"}, &g_brush_block))
        return -1;
    glPixelStorei(GL_PACK_ALIGNMENT, 1);
    glPixelStorei(GL_UNPACK_ALIGNMENT, 1);
    return gl_ok("init") ? 0 : -1;
}

// Runs `nsteps` reference steps on a W x H state in the reference texture
// layout.  rain (nullable) = {cx, cy, intensity, radius_sq}: the brush
// dispatched before every step.  Outputs may be null.  *seconds receives
// the wall time of the step loop (glFinish on both sides).
int glref_run(int W, int H, const float* terrain_rgba, const float* flow_rgba, float F, float K,
              float C, const float* rain, int nsteps, float* out_terrain, float* out_flow,
              uint16_t* out_flowmap, double* seconds) {
    if (!g_ctx) { g_err = "glref_init not called"; return -1; }
    // Any size: the dispatches below divide it with truncation, as
    // Terrain.cpp:167,258,264 do, so a tail of columns and rows that no work
    // group covers stays as uploaded.
    if (W <= 0 || H <= 0) {
        g_err = "grid must be at least 1 x 1";
        return -1;
    }
    GLuint terrain = make_tex(GL_RGBA32F, W, H);
    GLuint flow = make_tex(GL_RGBA32F, W, H);
    GLuint flowmap = make_tex(GL_RG16F, W, H);
    glBindTexture(GL_TEXTURE_2D, terrain);
    glTexSubImage2D(GL_TEXTURE_2D, 0, 0, 0, W, H, GL_RGBA, GL_FLOAT, terrain_rgba);
    glBindTexture(GL_TEXTURE_2D, flow);
    if (flow_rgba) {
        glTexSubImage2D(GL_TEXTURE_2D, 0, 0, 0, W, H, GL_RGBA, GL_FLOAT, flow_rgba);
    } else {
        std::vector<float> zero((size_t)W * H * 4, 0.0f);
        glTexSubImage2D(GL_TEXTURE_2D, 0, 0, 0, W, H, GL_RGBA, GL_FLOAT, zero.data());
    }
    {
        std::vector<uint16_t> zero((size_t)W * H * 2, 0);
        glBindTexture(GL_TEXTURE_2D, flowmap);
        glTexSubImage2D(GL_TEXTURE_2D, 0, 0, 0, W, H, GL_RG, GL_HALF_FLOAT, zero.data());
    }
    upload_block(g_sim_block, {F, K, C}, 5);
    if (rain) {
        // vec2 BrushPositionTexelCor occupies two floats at its offset
        std::vector<unsigned char> data((size_t)g_brush_block.size, 0);
        memcpy(&data[(size_t)g_brush_block.offs[0]], rain, 8);
        memcpy(&data[(size_t)g_brush_block.offs[1]], rain + 2, 4);
        memcpy(&data[(size_t)g_brush_block.offs[2]], rain + 3, 4);
        glBindBuffer(GL_UNIFORM_BUFFER, g_brush_block.buf);
        glBufferData(GL_UNIFORM_BUFFER, g_brush_block.size, data.data(), GL_DYNAMIC_DRAW);
        glBindBufferBase(GL_UNIFORM_BUFFER, 7, g_brush_block.buf);
    }
    if (!gl_ok("setup")) return -1;
    glFinish();
    auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < nsteps; ++s) {
        if (rain) {
            glBindImageTexture(0, terrain, 0, GL_FALSE, 0, GL_READ_WRITE, GL_RGBA32F);
            glUseProgram(g_brush);
            glDispatchCompute(W / 32, H / 32, 1);
            glMemoryBarrier(GL_SHADER_IMAGE_ACCESS_BARRIER_BIT);
        }
        glBindImageTexture(0, terrain, 0, GL_FALSE, 0, GL_READ_ONLY, GL_RGBA32F);
        glBindImageTexture(1, flow, 0, GL_FALSE, 0, GL_READ_WRITE, GL_RGBA32F);
        glUseProgram(g_update);
        glDispatchCompute(W / 16, H / 16, 1);
        glMemoryBarrier(GL_SHADER_IMAGE_ACCESS_BARRIER_BIT);
        glBindImageTexture(0, terrain, 0, GL_FALSE, 0, GL_READ_WRITE, GL_RGBA32F);
        glBindImageTexture(1, flow, 0, GL_FALSE, 0, GL_READ_ONLY, GL_RGBA32F);
        glBindImageTexture(2, flowmap, 0, GL_FALSE, 0, GL_WRITE_ONLY, GL_RG16F);
        glUseProgram(g_apply);
        glDispatchCompute(W / 16, H / 16, 1);
        glMemoryBarrier(GL_SHADER_IMAGE_ACCESS_BARRIER_BIT);
    }
    glFinish();
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (!gl_ok("steps")) return -1;
    glMemoryBarrier(GL_TEXTURE_UPDATE_BARRIER_BIT);
    if (out_terrain) {
        glBindTexture(GL_TEXTURE_2D, terrain);
        glGetTexImage(GL_TEXTURE_2D, 0, GL_RGBA, GL_FLOAT, out_terrain);
    }
    if (out_flow) {
        glBindTexture(GL_TEXTURE_2D, flow);
        glGetTexImage(GL_TEXTURE_2D, 0, GL_RGBA, GL_FLOAT, out_flow);
    }
    if (out_flowmap) {
        glBindTexture(GL_TEXTURE_2D, flowmap);
        glGetTexImage(GL_TEXTURE_2D, 0, GL_RG, GL_HALF_FLOAT, out_flowmap);
    }
    GLuint texs[3] = {terrain, flow, flowmap};
    glDeleteTextures(3, texs);
    return gl_ok("readback") ? 0 : -1;
}

// The mip chain of an RGBA32F texture as the reference builds it after each
// frame with steps: storage for `levels` levels (glTexStorage2D,
// Texture2D.cpp:12-21; a full chain is floor(log2(max(W,H))) + 1 levels,
// Texture.cpp:28-44), level 0 uploaded, then glGenerateMipmap
// (Texture2D::GenMipMaps, Texture2D.cpp:64-68).  Levels 1..levels-1 are read
// back one after the other into `out` (level l is max(1, W>>l) x
// max(1, H>>l) RGBA texels).  *seconds: wall time of glGenerateMipmap.
int glref_genmips(int W, int H, int levels, const float* rgba, float* out, double* seconds) {
    if (!g_ctx) { g_err = "glref_init not called"; return -1; }
    if (W <= 0 || H <= 0 || levels < 2) { g_err = "genmips: bad size or level count"; return -1; }
    GLuint t = 0;
    glGenTextures(1, &t);
    glBindTexture(GL_TEXTURE_2D, t);
    glTexStorage2D(GL_TEXTURE_2D, levels, GL_RGBA32F, W, H);
    glTexSubImage2D(GL_TEXTURE_2D, 0, 0, 0, W, H, GL_RGBA, GL_FLOAT, rgba);
    if (!gl_ok("genmips: setup")) return -1;
    glFinish();
    auto t0 = std::chrono::steady_clock::now();
    glGenerateMipmap(GL_TEXTURE_2D);
    glFinish();
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (!gl_ok("genmips: generate")) return -1;
    size_t off = 0;
    for (int l = 1; l < levels; ++l) {
        const int w = W >> l > 0 ? W >> l : 1, h = H >> l > 0 ? H >> l : 1;
        glGetTexImage(GL_TEXTURE_2D, l, GL_RGBA, GL_FLOAT, out + off);
        off += (size_t)w * h * 4;
    }
    glDeleteTextures(1, &t);
    return gl_ok("genmips: readback") ? 0 : -1;
}

}  // extern "C"
