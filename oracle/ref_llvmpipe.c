/*
 * ref_llvmpipe.c -- the reference's own GLSL compute shaders (depthToVert,
 * vertToNorm, track, reduce, integrate, raycast, trackSDF, reduceSDF, mip)
 * executed by Mesa's llvmpipe
 * on the host cores, headless: the swrast DRI driver is loaded through its
 * loader interface (GL/internal/dri_interface.h) and a surfaceless OpenGL 4.3
 * core context is made current, so no display, EGL or OSMesa is needed.
 *
 * Test infrastructure only (like the C oracle): the shader sources are
 * compiled into oracle/_ref/libkfref.so by oracle/build_ref.py from the
 * reference tree; nothing in the product loads this library.  The host side
 * below restates what gFusion.cpp does around each dispatch (textures and
 * formats gFusion.cpp:200-305, depthToVertex :496-526, vertexToNormal
 * :528-547, track :742-777, reduce :779-809, trackSDF :1040-1098, reduceSDF
 * :1100-1125, integrate :1234-1287, raycast :1289-1366, intensityProjection
 * :1368-1409): the same texture formats, image bindings, uniforms and work
 * group counts.  Matrices are 16 floats, column-major, as glm passes them.
 *
 * One deliberate difference: gFusion.cpp:1056, :1254, :1312 and :1372 bind
 * the 3D volume with layered = GL_FALSE.  By the GL specification that binds
 * layer 0 alone as a 2D image, so an iimage3D access is a type mismatch
 * (Mesa drops the loads and stores); NVIDIA's driver binds the whole volume,
 * which is what the reference was written and run against.  Here the volume
 * is bound layered, the behaviour the reference relies on.
 */
#include <dlfcn.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <GL/gl.h>
#include <GL/internal/dri_interface.h>

#define KFR_EXPORT __attribute__((visibility("default")))

/* shader sources, generated from the reference tree by build_ref.py */
extern const char *kfr_src_depthToVert, *kfr_src_vertToNorm, *kfr_src_track, *kfr_src_reduce,
    *kfr_src_integrate, *kfr_src_raycast, *kfr_src_trackSDF, *kfr_src_reduceSDF, *kfr_src_mip;

/* ---- GL entry points through the shared dispatch of libglapi ------------- */
typedef void *(*getproc_t)(const char *);
static getproc_t gp;
#define GLFN(ret, name, args) typedef ret(*PFN_##name) args; static PFN_##name p_##name;
GLFN(const GLubyte *, glGetString, (GLenum))
GLFN(GLuint, glCreateShader, (GLenum))
GLFN(void, glShaderSource, (GLuint, GLsizei, const char *const *, const GLint *))
GLFN(void, glCompileShader, (GLuint))
GLFN(void, glGetShaderiv, (GLuint, GLenum, GLint *))
GLFN(void, glGetShaderInfoLog, (GLuint, GLsizei, GLsizei *, char *))
GLFN(GLuint, glCreateProgram, (void))
GLFN(void, glAttachShader, (GLuint, GLuint))
GLFN(void, glLinkProgram, (GLuint))
GLFN(void, glGetProgramiv, (GLuint, GLenum, GLint *))
GLFN(void, glGetProgramInfoLog, (GLuint, GLsizei, GLsizei *, char *))
GLFN(void, glUseProgram, (GLuint))
GLFN(GLint, glGetUniformLocation, (GLuint, const char *))
GLFN(void, glUniformMatrix4fv, (GLint, GLsizei, GLboolean, const GLfloat *))
GLFN(void, glUniform4fv, (GLint, GLsizei, const GLfloat *))
GLFN(void, glUniform3fv, (GLint, GLsizei, const GLfloat *))
GLFN(void, glUniform2iv, (GLint, GLsizei, const GLint *))
GLFN(void, glUniform1f, (GLint, GLfloat))
GLFN(void, glGenTextures, (GLsizei, GLuint *))
GLFN(void, glDeleteTextures, (GLsizei, const GLuint *))
GLFN(void, glBindTexture, (GLenum, GLuint))
GLFN(void, glActiveTexture, (GLenum))
GLFN(void, glTexParameteri, (GLenum, GLenum, GLint))
GLFN(void, glTexStorage2D, (GLenum, GLsizei, GLenum, GLsizei, GLsizei))
GLFN(void, glTexStorage3D, (GLenum, GLsizei, GLenum, GLsizei, GLsizei, GLsizei))
GLFN(void, glTexSubImage2D, (GLenum, GLint, GLint, GLint, GLsizei, GLsizei, GLenum, GLenum, const void *))
GLFN(void, glTexSubImage3D,
     (GLenum, GLint, GLint, GLint, GLint, GLsizei, GLsizei, GLsizei, GLenum, GLenum, const void *))
GLFN(void, glGetTexImage, (GLenum, GLint, GLenum, GLenum, void *))
GLFN(void, glGenerateMipmap, (GLenum))
GLFN(void, glPixelStorei, (GLenum, GLint))
GLFN(void, glBindImageTexture, (GLuint, GLuint, GLint, GLboolean, GLint, GLenum, GLenum))
GLFN(void, glGenBuffers, (GLsizei, GLuint *))
GLFN(void, glDeleteBuffers, (GLsizei, const GLuint *))
GLFN(void, glBindBuffer, (GLenum, GLuint))
GLFN(void, glBufferData, (GLenum, long, const void *, GLenum))
GLFN(void, glBufferSubData, (GLenum, long, long, const void *))
GLFN(void, glGetBufferSubData, (GLenum, long, long, void *))
GLFN(void, glBindBufferBase, (GLenum, GLuint, GLuint))
GLFN(void, glDispatchCompute, (GLuint, GLuint, GLuint))
GLFN(void, glMemoryBarrier, (GLbitfield))
GLFN(void, glFinish, (void))
GLFN(GLenum, glGetError, (void))

/* enums of GL 4.3 core not in the GL 1.x header */
#define K_COMPUTE_SHADER 0x91B9
#define K_COMPILE_STATUS 0x8B81
#define K_LINK_STATUS 0x8B82
#define K_SSBO 0x90D2
#define K_DYNAMIC_COPY 0x88EA
#define K_TEXTURE_3D 0x806F
#define K_TEXTURE0 0x84C0
#define K_CLAMP_TO_EDGE 0x812F
#define K_R32F 0x822E
#define K_RGBA32F 0x8814
#define K_RG16I 0x8239
#define K_RG_INTEGER 0x8228
#define K_READ_WRITE 0x88BA
#define K_WRITE_ONLY 0x88B9
#define K_READ_ONLY 0x88B8
#define K_ALL_BARRIERS 0xFFFFFFFFu

static char g_err[2048];
static int fail(const char *fmt, const char *a)
{
    snprintf(g_err, sizeof(g_err), fmt, a ? a : "");
    return -1;
}
KFR_EXPORT const char *kfr_last_error(void) { return g_err; }

/* ---- headless context ------------------------------------------------------ */
static void lp_getDrawableInfo(__DRIdrawable *d, int *x, int *y, int *w, int *h, void *lp)
{
    (void)d; (void)lp;
    *x = *y = 0;
    *w = *h = 1;
}
static void lp_putImage(__DRIdrawable *d, int op, int x, int y, int w, int h, char *data, void *lp)
{
    (void)d; (void)op; (void)x; (void)y; (void)w; (void)h; (void)data; (void)lp;
}
static void lp_getImage(__DRIdrawable *d, int x, int y, int w, int h, char *data, void *lp)
{
    (void)d; (void)x; (void)y; (void)w; (void)h; (void)data; (void)lp;
}
static const __DRIswrastLoaderExtension g_loader = {{__DRI_SWRAST_LOADER, 1}, lp_getDrawableInfo, lp_putImage,
                                                     lp_getImage};
static const __DRIextension *g_loader_exts[] = {&g_loader.base, NULL};

static int g_ready = 0;
static char g_info[512];

enum { P_D2V, P_V2N, P_TRACK, P_REDUCE, P_INTEGRATE, P_RAYCAST, P_TSDF, P_RSDF, P_MIP, P_N };
static GLuint g_prog[P_N];

static int compile(const char *src, GLuint *prog)
{
    GLuint sh = p_glCreateShader(K_COMPUTE_SHADER);
    p_glShaderSource(sh, 1, &src, NULL);
    p_glCompileShader(sh);
    GLint ok = 0;
    p_glGetShaderiv(sh, K_COMPILE_STATUS, &ok);
    if (!ok) {
        char log[1536];
        p_glGetShaderInfoLog(sh, sizeof(log), NULL, log);
        return fail("shader compile: %s", log);
    }
    GLuint pr = p_glCreateProgram();
    p_glAttachShader(pr, sh);
    p_glLinkProgram(pr);
    p_glGetProgramiv(pr, K_LINK_STATUS, &ok);
    if (!ok) {
        char log[1536];
        p_glGetProgramInfoLog(pr, sizeof(log), NULL, log);
        return fail("program link: %s", log);
    }
    *prog = pr;
    return 0;
}

#define LOAD(name)                                           \
    do {                                                     \
        p_##name = (PFN_##name)gp(#name);                    \
        if (!p_##name) return fail("no GL entry %s", #name); \
    } while (0)

/* Creates the context and compiles the six programs (once per process). */
KFR_EXPORT int kfr_init(void)
{
    if (g_ready) return 0;
    void *g = dlopen("libglapi.so.0", RTLD_NOW | RTLD_GLOBAL);
    if (!g) return fail("libglapi: %s", dlerror());
    gp = (getproc_t)dlsym(g, "_glapi_get_proc_address");
    if (!gp) return fail("%s", "_glapi_get_proc_address missing");
    const char *drv = getenv("KFR_SWRAST_DRI");
    void *h = dlopen(drv ? drv : "/usr/lib/x86_64-linux-gnu/dri/swrast_dri.so", RTLD_NOW | RTLD_LOCAL);
    if (!h) return fail("swrast driver: %s", dlerror());
    const __DRIextension **(*get)(void) =
        (const __DRIextension **(*)(void))dlsym(h, "__driDriverGetExtensions_swrast");
    if (!get) return fail("%s", "__driDriverGetExtensions_swrast missing");
    const __DRIextension **ext = get();
    const __DRIcoreExtension *core = NULL;
    const __DRIswrastExtension *sw = NULL;
    for (int i = 0; ext[i]; ++i) {
        if (!strcmp(ext[i]->name, __DRI_CORE)) core = (const __DRIcoreExtension *)ext[i];
        if (!strcmp(ext[i]->name, __DRI_SWRAST)) sw = (const __DRIswrastExtension *)ext[i];
    }
    if (!core || !sw || sw->base.version < 4) return fail("%s", "driver lacks DRI_Core / DRI_SWRast v4");
    const __DRIconfig **configs = NULL;
    __DRIscreen *scr = sw->createNewScreen2(0, g_loader_exts, ext, &configs, NULL);
    if (!scr) return fail("%s", "createNewScreen2 failed");
    const uint32_t attribs[] = {__DRI_CTX_ATTRIB_MAJOR_VERSION, 4, __DRI_CTX_ATTRIB_MINOR_VERSION, 3};
    unsigned err = 0;
    __DRIcontext *ctx = sw->createContextAttribs(scr, __DRI_API_OPENGL_CORE, NULL, NULL, 2, attribs, &err, NULL);
    if (!ctx) return fail("%s", "createContextAttribs failed");
    if (!core->bindContext(ctx, NULL, NULL)) return fail("%s", "bindContext (surfaceless) failed");
    LOAD(glGetString); LOAD(glCreateShader); LOAD(glShaderSource); LOAD(glCompileShader); LOAD(glGetShaderiv);
    LOAD(glGetShaderInfoLog); LOAD(glCreateProgram); LOAD(glAttachShader); LOAD(glLinkProgram);
    LOAD(glGetProgramiv); LOAD(glGetProgramInfoLog); LOAD(glUseProgram); LOAD(glGetUniformLocation);
    LOAD(glUniformMatrix4fv); LOAD(glUniform4fv); LOAD(glUniform3fv); LOAD(glUniform2iv); LOAD(glUniform1f);
    LOAD(glGenTextures); LOAD(glDeleteTextures); LOAD(glBindTexture); LOAD(glActiveTexture);
    LOAD(glTexParameteri); LOAD(glTexStorage2D); LOAD(glTexStorage3D); LOAD(glTexSubImage2D);
    LOAD(glTexSubImage3D); LOAD(glGetTexImage); LOAD(glGenerateMipmap); LOAD(glPixelStorei);
    LOAD(glBindImageTexture); LOAD(glGenBuffers); LOAD(glDeleteBuffers); LOAD(glBindBuffer); LOAD(glBufferData);
    LOAD(glBufferSubData); LOAD(glGetBufferSubData); LOAD(glBindBufferBase); LOAD(glDispatchCompute);
    LOAD(glMemoryBarrier); LOAD(glFinish); LOAD(glGetError);
    snprintf(g_info, sizeof(g_info), "%s | %s", (const char *)p_glGetString(GL_RENDERER),
             (const char *)p_glGetString(GL_VERSION));
    const char *srcs[P_N] = {kfr_src_depthToVert, kfr_src_vertToNorm, kfr_src_track,
                             kfr_src_reduce,      kfr_src_integrate,  kfr_src_raycast,
                             kfr_src_trackSDF,    kfr_src_reduceSDF,  kfr_src_mip};
    for (int i = 0; i < P_N; ++i)
        if (compile(srcs[i], &g_prog[i])) return -1;
    p_glPixelStorei(GL_PACK_ALIGNMENT, 1);
    p_glPixelStorei(GL_UNPACK_ALIGNMENT, 1);
    g_ready = 1;
    return 0;
}

KFR_EXPORT const char *kfr_info(void) { return g_info; }

/* ---- gFusion's objects (initTextures / initVolume / allocateBuffers) ------- */
static int g_w = 0, g_h = 0, g_n[3] = {0, 0, 0};
static GLuint t_depth, t_vert, t_norm, t_refV, t_refN, t_diff, t_trackim, t_test, t_mip, t_vol, b_red, b_out,
    b_sdf, b_sdfout;

static GLuint make_tex(GLenum target, int levels, int w, int h, int d, GLenum fmt)
{
    /* createTexture, gFusion.cpp:171-198 */
    GLuint t;
    p_glGenTextures(1, &t);
    p_glBindTexture(target, t);
    p_glTexParameteri(target, GL_TEXTURE_MAG_FILTER, GL_LINEAR);
    p_glTexParameteri(target, GL_TEXTURE_MIN_FILTER, GL_LINEAR);
    p_glTexParameteri(target, GL_TEXTURE_WRAP_S, K_CLAMP_TO_EDGE);
    p_glTexParameteri(target, GL_TEXTURE_WRAP_T, K_CLAMP_TO_EDGE);
    if (target == GL_TEXTURE_2D) p_glTexStorage2D(target, levels, fmt, w, h);
    else p_glTexStorage3D(target, levels, fmt, w, h, d);
    return t;
}

/* Depth size w x h and volume nx x ny x nz (re-created when they change). */
KFR_EXPORT int kfr_setup(int w, int h, int nx, int ny, int nz)
{
    if (!g_ready && kfr_init()) return -1;
    if (w == g_w && h == g_h && nx == g_n[0] && ny == g_n[1] && nz == g_n[2]) return 0;
    if (g_w) {
        GLuint ts[10] = {t_depth, t_vert, t_norm, t_refV, t_refN, t_diff, t_trackim, t_test, t_mip, t_vol};
        p_glDeleteTextures(10, ts);
        GLuint bs[4] = {b_red, b_out, b_sdf, b_sdfout};
        p_glDeleteBuffers(4, bs);
    }
    t_depth = make_tex(GL_TEXTURE_2D, 3, w, h, 1, K_R32F);
    t_vert = make_tex(GL_TEXTURE_2D, 3, w, h, 1, K_RGBA32F);
    t_norm = make_tex(GL_TEXTURE_2D, 3, w, h, 1, K_RGBA32F);
    t_refV = make_tex(GL_TEXTURE_2D, 1, w, h, 1, K_RGBA32F);
    t_refN = make_tex(GL_TEXTURE_2D, 1, w, h, 1, K_RGBA32F);
    t_diff = make_tex(GL_TEXTURE_2D, 1, w, h, 1, K_R32F);
    t_trackim = make_tex(GL_TEXTURE_2D, 1, w, h, 1, K_RGBA32F);
    t_test = make_tex(GL_TEXTURE_2D, 1, w, h, 1, K_RGBA32F);
    t_mip = make_tex(GL_TEXTURE_2D, 1, w, h, 1, K_RGBA32F);
    t_vol = make_tex(K_TEXTURE_3D, 1, nx, ny, nz, K_RG16I);
    /* the reduction buffer holds one 32-byte track record per level-0 pixel
       (allocateBuffers, gFusion.cpp:297-305); depthToVert's dead write of
       Position3D lands in it too */
    p_glGenBuffers(1, &b_red);
    p_glBindBuffer(K_SSBO, b_red);
    p_glBufferData(K_SSBO, (long)w * h * 32, NULL, K_DYNAMIC_COPY);
    p_glGenBuffers(1, &b_out);
    p_glBindBuffer(K_SSBO, b_out);
    p_glBufferData(K_SSBO, 256 * 4, NULL, K_DYNAMIC_COPY);
    /* point-to-volume records {h, D, J[6]} per level-0 pixel and their
       8 x 32 partials (gFusion.cpp:345-353) */
    p_glGenBuffers(1, &b_sdf);
    p_glBindBuffer(K_SSBO, b_sdf);
    p_glBufferData(K_SSBO, (long)w * h * 32, NULL, K_DYNAMIC_COPY);
    p_glGenBuffers(1, &b_sdfout);
    p_glBindBuffer(K_SSBO, b_sdfout);
    p_glBufferData(K_SSBO, 256 * 4, NULL, K_DYNAMIC_COPY);
    p_glBindBufferBase(K_SSBO, 0, b_red);
    p_glBindBufferBase(K_SSBO, 1, b_out);
    p_glBindBufferBase(K_SSBO, 12, b_sdf);
    p_glBindBufferBase(K_SSBO, 13, b_sdfout);
    g_w = w;
    g_h = h;
    g_n[0] = nx;
    g_n[1] = ny;
    g_n[2] = nz;
    if (p_glGetError() != GL_NO_ERROR) return fail("%s", "GL error in setup");
    return 0;
}

static int divup(int a, int b) { return (a + b - 1) / b; }
static GLint U(int p, const char *name) { return p_glGetUniformLocation(g_prog[p], name); }

/* getInverseCameraMatrix (gFusion.h:451-459), column-major */
static void inv_cam(const float c[4], float m[16])
{
    memset(m, 0, 64);
    m[0] = 1.0f / c[0];
    m[5] = 1.0f / c[1];
    m[8] = -c[2] / c[0];
    m[9] = -c[3] / c[1];
    m[10] = 1.0f;
    m[15] = 1.0f;
}

/* depthToVertex(float *depthArray), gFusion.cpp:496-526 */
KFR_EXPORT int kfr_depth_to_vertex(const float *depth_mm, const float cam[4])
{
    p_glActiveTexture(K_TEXTURE0);
    p_glBindTexture(GL_TEXTURE_2D, t_depth);
    p_glTexSubImage2D(GL_TEXTURE_2D, 0, 0, 0, g_w, g_h, GL_RED, GL_FLOAT, depth_mm);
    p_glGenerateMipmap(GL_TEXTURE_2D);
    p_glUseProgram(g_prog[P_D2V]);
    p_glBindBufferBase(K_SSBO, 0, b_red);
    for (int i = 0; i < 3; ++i) {
        const float s = (float)(1 << i);
        const float cl[4] = {cam[0] / s, cam[1] / s, cam[2] / s, cam[3] / s};
        float iK[16];
        inv_cam(cl, iK);
        p_glBindImageTexture(0, t_depth, i, GL_FALSE, 0, K_READ_WRITE, K_R32F);
        p_glBindImageTexture(1, t_vert, i, GL_FALSE, 0, K_READ_WRITE, K_RGBA32F);
        p_glUniformMatrix4fv(U(P_D2V, "invK"), 1, GL_FALSE, iK);
        p_glUniform4fv(U(P_D2V, "camPams"), 1, cam);
        p_glDispatchCompute(divup(g_w >> i, 32), divup(g_h >> i, 32), 1);
        p_glMemoryBarrier(K_ALL_BARRIERS);
    }
    return p_glGetError() == GL_NO_ERROR ? 0 : fail("%s", "GL error in depthToVertex");
}

/* vertexToNormal(), gFusion.cpp:528-547 */
KFR_EXPORT int kfr_vertex_to_normal(void)
{
    p_glUseProgram(g_prog[P_V2N]);
    for (int i = 0; i < 3; ++i) {
        p_glBindImageTexture(0, t_vert, i, GL_FALSE, 0, K_READ_WRITE, K_RGBA32F);
        p_glBindImageTexture(1, t_norm, i, GL_FALSE, 0, K_READ_WRITE, K_RGBA32F);
        p_glDispatchCompute(divup(g_w >> i, 32), divup(g_h >> i, 32), 1);
        p_glMemoryBarrier(K_ALL_BARRIERS);
    }
    return p_glGetError() == GL_NO_ERROR ? 0 : fail("%s", "GL error in vertexToNormal");
}

/* which: 0 depth (float mm), 1 vertex, 2 normal (float4) at `level`;
   3 reference vertex, 4 reference normal (level 0) */
static GLuint map_tex(int which, GLenum *fmt, int *comps)
{
    *fmt = which == 0 ? GL_RED : GL_RGBA;
    *comps = which == 0 ? 1 : 4;
    switch (which) {
    case 0: return t_depth;
    case 1: return t_vert;
    case 2: return t_norm;
    case 3: return t_refV;
    case 4: return t_refN;
    }
    return 0;
}

KFR_EXPORT int kfr_read_map(int which, int level, float *out)
{
    GLenum fmt;
    int comps;
    GLuint t = map_tex(which, &fmt, &comps);
    if (!t) return fail("%s", "unknown map");
    p_glBindTexture(GL_TEXTURE_2D, t);
    p_glGetTexImage(GL_TEXTURE_2D, level, fmt, GL_FLOAT, out);
    return p_glGetError() == GL_NO_ERROR ? 0 : fail("%s", "GL error in read_map");
}

KFR_EXPORT int kfr_write_map(int which, int level, const float *in)
{
    GLenum fmt;
    int comps;
    GLuint t = map_tex(which, &fmt, &comps);
    if (!t) return fail("%s", "unknown map");
    p_glBindTexture(GL_TEXTURE_2D, t);
    p_glTexSubImage2D(GL_TEXTURE_2D, level, 0, 0, g_w >> level, g_h >> level, fmt, GL_FLOAT, in);
    return p_glGetError() == GL_NO_ERROR ? 0 : fail("%s", "GL error in write_map");
}

/* track(layer), gFusion.cpp:742-777: view = K * inverse(m_pose), Ttrack = m_pose */
KFR_EXPORT int kfr_track(int layer, const float view[16], const float Ttrack[16], float dist_th, float normal_th)
{
    p_glUseProgram(g_prog[P_TRACK]);
    p_glUniformMatrix4fv(U(P_TRACK, "view"), 1, GL_FALSE, view);
    p_glUniformMatrix4fv(U(P_TRACK, "Ttrack"), 1, GL_FALSE, Ttrack);
    p_glUniform1f(U(P_TRACK, "dist_threshold"), dist_th);
    p_glUniform1f(U(P_TRACK, "normal_threshold"), normal_th);
    p_glBindImageTexture(0, t_vert, layer, GL_FALSE, 0, K_READ_WRITE, K_RGBA32F);
    p_glBindImageTexture(1, t_norm, layer, GL_FALSE, 0, K_READ_WRITE, K_RGBA32F);
    p_glBindImageTexture(2, t_refV, 0, GL_FALSE, 0, K_READ_WRITE, K_RGBA32F);
    p_glBindImageTexture(3, t_refN, 0, GL_FALSE, 0, K_READ_WRITE, K_RGBA32F);
    p_glBindImageTexture(4, t_diff, 0, GL_FALSE, 0, K_READ_WRITE, K_R32F);
    p_glBindImageTexture(5, t_trackim, 0, GL_FALSE, 0, K_READ_WRITE, K_RGBA32F);
    p_glBindBufferBase(K_SSBO, 0, b_red);
    p_glDispatchCompute(divup(g_w >> layer, 32), divup(g_h >> layer, 32), 1);
    p_glMemoryBarrier(K_ALL_BARRIERS);
    return p_glGetError() == GL_NO_ERROR ? 0 : fail("%s", "GL error in track");
}

/* The track records of the first n pixels ({int result; float error;
   float J[6]}, track.cs:18-27) as planes. */
KFR_EXPORT int kfr_read_track(int n, int32_t *result, float *error, float *J)
{
    float *rec = (float *)malloc((size_t)n * 32);
    if (!rec) return fail("%s", "out of memory");
    p_glBindBuffer(K_SSBO, b_red);
    p_glGetBufferSubData(K_SSBO, 0, (long)n * 32, rec);
    for (int i = 0; i < n; ++i) {
        memcpy(&result[i], &rec[8 * i], 4);
        error[i] = rec[8 * i + 1];
        memcpy(&J[6 * i], &rec[8 * i + 2], 24);
    }
    free(rec);
    return p_glGetError() == GL_NO_ERROR ? 0 : fail("%s", "GL error in read_track");
}

/* Writes track records (the reduce input) for the first n pixels. */
KFR_EXPORT int kfr_write_track(int n, const int32_t *result, const float *error, const float *J)
{
    float *rec = (float *)calloc((size_t)n, 32);
    if (!rec) return fail("%s", "out of memory");
    for (int i = 0; i < n; ++i) {
        memcpy(&rec[8 * i], &result[i], 4);
        rec[8 * i + 1] = error[i];
        memcpy(&rec[8 * i + 2], &J[6 * i], 24);
    }
    p_glBindBuffer(K_SSBO, b_red);
    p_glBufferSubData(K_SSBO, 0, (long)n * 32, rec);
    free(rec);
    return p_glGetError() == GL_NO_ERROR ? 0 : fail("%s", "GL error in write_track");
}

/* reduce(layer), gFusion.cpp:779-809: 8 work groups of 112; out = 8 x 32 */
KFR_EXPORT int kfr_reduce(int layer, float out[256])
{
    const GLint size[2] = {g_w >> layer, g_h >> layer};
    p_glUseProgram(g_prog[P_REDUCE]);
    p_glBindBufferBase(K_SSBO, 0, b_red);
    p_glBindBufferBase(K_SSBO, 1, b_out);
    p_glUniform2iv(U(P_REDUCE, "imageSize"), 1, size);
    p_glDispatchCompute(8, 1, 1);
    p_glMemoryBarrier(K_ALL_BARRIERS);
    p_glBindBuffer(K_SSBO, b_out);
    p_glGetBufferSubData(K_SSBO, 0, 256 * 4, out);
    return p_glGetError() == GL_NO_ERROR ? 0 : fail("%s", "GL error in reduce");
}

/* The RG16I volume, interleaved (value, weight) int16 pairs, x fastest. */
KFR_EXPORT int kfr_write_volume(const int16_t *vol)
{
    p_glBindTexture(K_TEXTURE_3D, t_vol);
    p_glTexSubImage3D(K_TEXTURE_3D, 0, 0, 0, 0, g_n[0], g_n[1], g_n[2], K_RG_INTEGER, GL_SHORT, vol);
    return p_glGetError() == GL_NO_ERROR ? 0 : fail("%s", "GL error in write_volume");
}

KFR_EXPORT int kfr_read_volume(int16_t *vol)
{
    p_glBindTexture(K_TEXTURE_3D, t_vol);
    p_glGetTexImage(K_TEXTURE_3D, 0, K_RG_INTEGER, GL_SHORT, vol);
    return p_glGetError() == GL_NO_ERROR ? 0 : fail("%s", "GL error in read_volume");
}

/* integrate(), gFusion.cpp:1234-1287: invTrack = inverse(m_pose), K =
   getCameraMatrix(camPams); the depth texture's level 0 is the input */
KFR_EXPORT int kfr_integrate(const float invTrack[16], const float K[16], float mu, float max_weight,
                             const float vol_dim[3])
{
    const float vsize[3] = {(float)g_n[0], (float)g_n[1], (float)g_n[2]};
    p_glUseProgram(g_prog[P_INTEGRATE]);
    p_glUniformMatrix4fv(U(P_INTEGRATE, "invTrack"), 1, GL_FALSE, invTrack);
    p_glUniformMatrix4fv(U(P_INTEGRATE, "K"), 1, GL_FALSE, K);
    p_glUniform1f(U(P_INTEGRATE, "mu"), mu);
    p_glUniform1f(U(P_INTEGRATE, "maxWeight"), max_weight);
    p_glUniform3fv(U(P_INTEGRATE, "volDim"), 1, vol_dim);
    p_glUniform3fv(U(P_INTEGRATE, "volSize"), 1, vsize);
    p_glBindImageTexture(0, t_vol, 0, GL_TRUE, 0, K_READ_WRITE, K_RG16I);
    p_glBindImageTexture(1, t_depth, 0, GL_FALSE, 0, K_READ_WRITE, K_R32F);
    p_glDispatchCompute(divup(g_n[0], 32), divup(g_n[1], 32), 1);
    p_glMemoryBarrier(K_ALL_BARRIERS);
    return p_glGetError() == GL_NO_ERROR ? 0 : fail("%s", "GL error in integrate");
}

/* raycast(), gFusion.cpp:1289-1366: view = m_pose * invK */
KFR_EXPORT int kfr_raycast(const float view[16], float near_plane, float far_plane, float step, float large_step,
                           const float vol_dim[3])
{
    const float vsize[3] = {(float)g_n[0], (float)g_n[1], (float)g_n[2]};
    p_glUseProgram(g_prog[P_RAYCAST]);
    p_glUniformMatrix4fv(U(P_RAYCAST, "view"), 1, GL_FALSE, view);
    p_glUniform1f(U(P_RAYCAST, "nearPlane"), near_plane);
    p_glUniform1f(U(P_RAYCAST, "farPlane"), far_plane);
    p_glUniform1f(U(P_RAYCAST, "step"), step);
    p_glUniform1f(U(P_RAYCAST, "largeStep"), large_step);
    p_glUniform3fv(U(P_RAYCAST, "volDim"), 1, vol_dim);
    p_glUniform3fv(U(P_RAYCAST, "volSize"), 1, vsize);
    p_glBindImageTexture(0, t_vol, 0, GL_TRUE, 0, K_READ_WRITE, K_RG16I);
    p_glBindImageTexture(1, t_refV, 0, GL_FALSE, 0, K_WRITE_ONLY, K_RGBA32F);
    p_glBindImageTexture(2, t_refN, 0, GL_FALSE, 0, K_WRITE_ONLY, K_RGBA32F);
    p_glActiveTexture(K_TEXTURE0);
    p_glBindTexture(K_TEXTURE_3D, t_vol);
    p_glDispatchCompute(divup(g_w, 32), divup(g_h, 32), 1);
    p_glMemoryBarrier(K_ALL_BARRIERS);
    return p_glGetError() == GL_NO_ERROR ? 0 : fail("%s", "GL error in raycast");
}

/* trackSDF(layer, camToWorld), gFusion.cpp:1040-1098: c = 0.02 * 0.1,
   eps = 10e-9; records go to the SSBO at binding 12 */
KFR_EXPORT int kfr_track_sdf(int layer, const float camToWorld[16], const float vol_dim[3])
{
    const float vsize[3] = {(float)g_n[0], (float)g_n[1], (float)g_n[2]};
    const GLint size[2] = {g_w >> layer, g_h >> layer};
    p_glUseProgram(g_prog[P_TSDF]);
    p_glActiveTexture(K_TEXTURE0);
    p_glBindTexture(K_TEXTURE_3D, t_vol);
    p_glBindImageTexture(0, t_vol, 0, GL_TRUE, 0, K_READ_WRITE, K_RG16I);
    p_glBindImageTexture(1, t_vert, layer, GL_FALSE, 0, K_READ_WRITE, K_RGBA32F);
    p_glBindImageTexture(2, t_test, 0, GL_FALSE, 0, K_WRITE_ONLY, K_RGBA32F);
    p_glUniformMatrix4fv(U(P_TSDF, "Ttrack"), 1, GL_FALSE, camToWorld);
    p_glUniform2iv(U(P_TSDF, "imageSize"), 1, size);
    p_glUniform1f(U(P_TSDF, "c"), 0.02f * 0.1f);
    p_glUniform1f(U(P_TSDF, "eps"), 10e-9f);
    p_glUniform3fv(U(P_TSDF, "volDim"), 1, vol_dim);
    p_glUniform3fv(U(P_TSDF, "volSize"), 1, vsize);
    p_glBindBufferBase(K_SSBO, 12, b_sdf);
    p_glDispatchCompute(divup(g_w >> layer, 32), divup(g_h >> layer, 32), 1);
    p_glMemoryBarrier(K_ALL_BARRIERS);
    return p_glGetError() == GL_NO_ERROR ? 0 : fail("%s", "GL error in trackSDF");
}

/* The first n point-to-volume records, 8 floats each ({h, D, J[6]}). */
KFR_EXPORT int kfr_read_sdf(int n, float *rec)
{
    p_glBindBuffer(K_SSBO, b_sdf);
    p_glGetBufferSubData(K_SSBO, 0, (long)n * 32, rec);
    return p_glGetError() == GL_NO_ERROR ? 0 : fail("%s", "GL error in read_sdf");
}

KFR_EXPORT int kfr_write_sdf(int n, const float *rec)
{
    p_glBindBuffer(K_SSBO, b_sdf);
    p_glBufferSubData(K_SSBO, 0, (long)n * 32, rec);
    return p_glGetError() == GL_NO_ERROR ? 0 : fail("%s", "GL error in write_sdf");
}

/* reduceSDF(layer), gFusion.cpp:1100-1125: 8 work groups of 112 */
KFR_EXPORT int kfr_reduce_sdf(int layer, float out[256])
{
    const GLint size[2] = {g_w >> layer, g_h >> layer};
    p_glUseProgram(g_prog[P_RSDF]);
    p_glBindBufferBase(K_SSBO, 12, b_sdf);
    p_glBindBufferBase(K_SSBO, 13, b_sdfout);
    p_glUniform2iv(U(P_RSDF, "imageSize"), 1, size);
    p_glDispatchCompute(8, 1, 1);
    p_glMemoryBarrier(K_ALL_BARRIERS);
    p_glBindBuffer(K_SSBO, b_sdfout);
    p_glGetBufferSubData(K_SSBO, 0, 256 * 4, out);
    return p_glGetError() == GL_NO_ERROR ? 0 : fail("%s", "GL error in reduceSDF");
}

/* intensityProjection(), gFusion.cpp:1368-1409 (mip.cs): view = m_pose * invK;
   the RGBA32F result image is returned in out (h x w x 4) */
KFR_EXPORT int kfr_intensity_projection(const float view[16], float near_plane, float far_plane, float step,
                                        float large_step, const float vol_dim[3], float *out)
{
    const float vsize[3] = {(float)g_n[0], (float)g_n[1], (float)g_n[2]};
    p_glUseProgram(g_prog[P_MIP]);
    p_glBindImageTexture(0, t_vol, 0, GL_TRUE, 0, K_READ_ONLY, K_RG16I);
    p_glBindImageTexture(1, t_mip, 0, GL_FALSE, 0, K_WRITE_ONLY, K_RGBA32F);
    p_glUniformMatrix4fv(U(P_MIP, "view"), 1, GL_FALSE, view);
    p_glUniform1f(U(P_MIP, "nearPlane"), near_plane);
    p_glUniform1f(U(P_MIP, "farPlane"), far_plane);
    p_glUniform1f(U(P_MIP, "step"), step);
    p_glUniform1f(U(P_MIP, "largeStep"), large_step);
    p_glUniform3fv(U(P_MIP, "volDim"), 1, vol_dim);
    p_glUniform3fv(U(P_MIP, "volSize"), 1, vsize);
    p_glDispatchCompute(divup(g_w, 32), divup(g_h, 32), 1);
    p_glMemoryBarrier(K_ALL_BARRIERS);
    p_glBindTexture(GL_TEXTURE_2D, t_mip);
    p_glGetTexImage(GL_TEXTURE_2D, 0, GL_RGBA, GL_FLOAT, out);
    return p_glGetError() == GL_NO_ERROR ? 0 : fail("%s", "GL error in intensityProjection");
}

KFR_EXPORT void kfr_finish(void) { p_glFinish(); }
