// puresoft3d.h -- C++ host facade with the reference's pipeline API
// (src/puresoft3d/pipeline.h:24-156, vbo.h, proc.h, defs.h) over the C ABI of
// include/psgpu.h.  Code written against PuresoftPipeline keeps its calls;
// the processors it adds must be the built-in ones, which have GPU twins.
// Errors are raised as the reference raises them: std::out_of_range,
// std::invalid_argument, std::bad_alloc, and std::runtime_error for device
// failures.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <utility>
#include <vector>

#include "../../include/psgpu.h"

// ---- defs.h -----------------------------------------------------------------
typedef struct {
    size_t capacity;
    void* data;
} PURESOFTUNIFORM;

typedef struct {
    union {
        struct { unsigned char bgra[4]; } ary;
        struct { unsigned char b, g, r, a; } elems;
        unsigned int i32;
    };
} PURESOFTBGRA;

typedef struct {
    unsigned int width;
    unsigned int scanline;
    unsigned int height;
    unsigned int elemLen;
    void* pixels;
} PURESOFTIMGBUFF32;

const PURESOFTBGRA PURESOFTBGRA_BLACK = {{{{0, 0, 0, 0}}}};

// ---- pipeline.h:17-20 ---------------------------------------------------------
const int BEHAVIOR_UPDATE_DEPTH = PSGPU_UPDATE_DEPTH;
const int BEHAVIOR_TEST_DEPTH = PSGPU_TEST_DEPTH;
const int BEHAVIOR_FACE_CULLING = PSGPU_FACE_CULLING;
const int BEHAVIOR_ALPHABLEND = PSGPU_ALPHABLEND;

// ---- fbo.h: a frame buffer in host memory ---------------------------------------
// The reference's FBO (fbo.h:6-100) with its constructor and buffer access:
// the frame and depth buffers a host post-processor is handed, and the
// targets setFBO binds.  Rows are `scanline` bytes apart; a top-down FBO
// holds display row y at memory row height-1-y (fbo.cpp:103-106).
class PuresoftFBO {
public:
    enum WRAPMODE { CLAMP = PSGPU_CLAMP, WRAP = PSGPU_WRAP };
    enum LAYER { LAYER_DEFAULT, LAYER_XPOS = LAYER_DEFAULT, LAYER_XNEG, LAYER_YPOS, LAYER_YNEG, LAYER_ZPOS, LAYER_ZNEG,
                 LAYER_MAX };

    PuresoftFBO(int width, int scanline, int height, int elemLen, bool topDown = false, void* externalBuffer = nullptr,
                WRAPMODE wrapMode = CLAMP, int extraLayers = 0);
    void setBuffer(void* externalBuffer = nullptr);   // null: back to the FBO's own memory
    void* getBuffer(void) { return m_buffer; }
    int getWidth(void) const { return m_width; }
    int getHeight(void) const { return m_height; }
    int getScanline(void) const { return m_scanline; }
    int getElemLen(void) const { return m_elemLen; }
    size_t getBytes(void) const { return (size_t)m_scanline * m_height; }
    bool isTopDown(void) const { return m_topDown; }
    WRAPMODE getWrapMode(void) const { return m_wrap; }
    int getExtraLayers(void) const { return m_extraLayers; }

private:
    int m_width, m_scanline, m_height, m_elemLen;
    bool m_topDown;
    WRAPMODE m_wrap;
    int m_extraLayers;
    std::vector<unsigned char> m_own;
    void* m_buffer;
};

// ---- vbo.h: a vertex stream; contents live on the host until attached ----------
class PuresoftVBO {
public:
    PuresoftVBO(size_t unitBytes, size_t unitCount);
    void updateContent(const void* src);
    size_t unitBytes() const { return m_unitBytes; }
    size_t unitCount() const { return m_unitCount; }

private:
    friend class PuresoftPipeline;
    size_t m_unitBytes, m_unitCount;
    std::vector<unsigned char> m_host;
    psgpu_ctx* m_ctx = nullptr;   // set once attached
    int m_handle = -1;
};

// ---- proc.h: processors are the built-in types with a GPU twin ------------------
class PuresoftProcessor {
public:
    explicit PuresoftProcessor(int kind) : m_kind(kind) {}
    virtual ~PuresoftProcessor() {}
    int gpuKind() const { return m_kind; }

private:
    int m_kind;
};
class PuresoftVertexProcessor : public PuresoftProcessor { public: using PuresoftProcessor::PuresoftProcessor; };
class PuresoftInterpolationProcessor : public PuresoftProcessor { public: using PuresoftProcessor::PuresoftProcessor; };
class PuresoftFragmentProcessor : public PuresoftProcessor { public: using PuresoftProcessor::PuresoftProcessor; };

#define PS_DECLARE_PROCS(NAME, KIND)                                                                           \
    class VertexProcesser##NAME : public PuresoftVertexProcessor { public: VertexProcesser##NAME() : PuresoftVertexProcessor(KIND) {} }; \
    class InterpolationProcessor##NAME : public PuresoftInterpolationProcessor { public: InterpolationProcessor##NAME() : PuresoftInterpolationProcessor(KIND) {} }; \
    class FragmentProcessor##NAME : public PuresoftFragmentProcessor { public: FragmentProcessor##NAME() : PuresoftFragmentProcessor(KIND) {} };

PS_DECLARE_PROCS(DEF01, PSGPU_PROC_DEF01)      // tex1light1.h
PS_DECLARE_PROCS(DEF02, PSGPU_PROC_DEF02)      // colr1light1.h
PS_DECLARE_PROCS(DEF03, PSGPU_PROC_DEF03)      // tex1bump1light1.h
PS_DECLARE_PROCS(DEF04, PSGPU_PROC_DEF04)      // skybox.h
PS_DECLARE_PROCS(DEF05, PSGPU_PROC_DEF05)      // shadow.h
PS_DECLARE_PROCS(LAMBERT, PSGPU_PROC_LAMBERT)  // Lambert x bilinear, directional light
PS_DECLARE_PROCS(BLINN, PSGPU_PROC_BLINN)      // Blinn-Phong + normal map, trilinear
PS_DECLARE_PROCS(SHADOWED, PSGPU_PROC_SHADOWED) // the same, lit through a projective shadow map (samplrproj.cpp)
PS_DECLARE_PROCS(CLOUD, PSGPU_PROC_CLOUD)      // the demo's cloud, blended (test/testproc.cpp:384-574)
PS_DECLARE_PROCS(CLOUD_SHADOW, PSGPU_PROC_CLOUD_SHADOW) // its shadow pass, discarding thin cloud (:589-686)
#undef PS_DECLARE_PROCS

// ---- proc.h:89-94: the post-processor ----------------------------------------------
// Two forms.  The reference's: process(threadIndex, threadCount, frame, depth)
// is host code over the frame (the colour target of FBO slot 0) and the depth
// buffer, called once per thread (post.cpp:3-18).  The pipeline finishes the
// draws before it, copies both buffers to host FBOs, calls it from
// threadCount threads and writes both buffers back.  The device form (no
// counterpart in the reference): process(frame, stream) gets the device
// addresses of the same buffers and the context stream and enqueues its own
// kernels or copies there, with no copies through host memory.  A subclass
// overrides one of the two; the device form's default returns kHostForm,
// which selects the host form.
class PuresoftPostProcessor {
public:
    enum { kHostForm = 1 };
    virtual ~PuresoftPostProcessor() {}
    virtual void process(int threadIndex, int threadCount, PuresoftFBO* frame, PuresoftFBO* depth) {}
    virtual int process(const psgpu_post_frame& frame, void* stream) { return kHostForm; }   // 0 or a negative status
};

// ---- rndr.h: a headless renderer owning the host copy of the display ------------
class PuresoftHeadlessRenderer {
public:
    PuresoftHeadlessRenderer(int w, int h) : m_w(w), m_h(h), m_pixels((size_t)w * h * 4) {}
    void getDesc(PURESOFTIMGBUFF32* d) { d->width = m_w; d->scanline = m_w * 4; d->height = m_h; d->elemLen = 4; d->pixels = m_pixels.data(); }
    void* buffer() { return m_pixels.data(); }

private:
    int m_w, m_h;
    std::vector<unsigned char> m_pixels;
};

// ---- pipeline.h ------------------------------------------------------------------
struct PsPostCall;   // a post-processor as the C ABI calls it back
class PuresoftPipeline {
public:
    PuresoftPipeline(uintptr_t canvasWindow, int deviceWidth, int deviceHeight, PuresoftHeadlessRenderer* rndr = nullptr,
                     int device = 0);
    ~PuresoftPipeline();

    // texture api
    int createTexture(const PURESOFTIMGBUFF32* image, int extraLayers = 0,
                      PuresoftFBO::WRAPMODE mode = PuresoftFBO::CLAMP, int mipLevels = 1);
    void destroyTexture(int idx);
    // tex.cpp:30-45: the layer's layout, and in image->pixels a writable
    // pointer to its texels.  The pointer is to a host mirror of the layer
    // that the pipeline keeps; what is written through it reaches the GPU
    // before the next draw, read or post-process (and stays valid until the
    // texture is destroyed).  Writes made after that need another getTexture.
    void getTexture(int idx, PURESOFTIMGBUFF32* image, PuresoftFBO::LAYER layer = PuresoftFBO::LAYER_DEFAULT);
    void setTexturePixels(int idx, const void* pixels, PuresoftFBO::LAYER layer = PuresoftFBO::LAYER_DEFAULT);
    void generateMipmaps(int idx);

    // processor api
    int addProcessor(PuresoftProcessor* proc);
    void destroyProcessor(int idx);
    int createProgramme(int vid, int iid, int fid);
    void destroyProgramme(int idx);
    void useProgramme(int idx);

    // vao api
    int createVAO(void);
    PuresoftVBO* attachVBO(int vao, int idx, PuresoftVBO* vbo);
    PuresoftVBO* detachVBO(int vao, int idx);
    PuresoftVBO* getVBO(int vao, int idx);
    void destroyVAO(int vao);

    // rendering api
    void setRenderer(PuresoftHeadlessRenderer* rndr);   // not owned; null: a renderer of the device size
    void setViewport(int width, int height, uintptr_t canvasWindow = 0);
    void setDepth(int idx = -1);
    // a texture of this pipeline as FBO slot idx (PSGPU_DISPLAY: the display for
    // slot 0, none for the others); no counterpart in the reference, whose
    // setFBO takes an FBO object (below)
    void setFBOTexture(int idx, int textureOrDisplay);
    // pipeline.cpp:260-275: a host FBO as target (null: the display for slot 0,
    // none for slots 1..15).  Slot 0's FBO is rendered on the GPU; like the
    // reference's synchronous drawVAO, each draw leaves the result in the FBO's
    // memory (its contents go to the GPU before the draw and come back after).
    void setFBO(int idx, PuresoftFBO* fbo);
    void setUniform(int idx, const void* data, size_t len);
    void drawVAO(int vao, bool callerThrdForFragProc = false);
    void swapBuffers(void);                            // display -> renderer buffer
    void postProcess(PuresoftPostProcessor* processor); // not owned; must outlive the frame
    void enable(int behavior);
    void disable(int behavior);
    void clearDepth(float furthest = 1.0f);
    void clearColour(PURESOFTBGRA bkgnd = PURESOFTBGRA_BLACK);

    // debug api (dbg.cpp): idx -1 is the current depth buffer, -2 the display
    void saveTexture(int idx, const char* path, bool dataIsFloat);
    void saveTexture(int idx, const wchar_t* path, bool dataIsFloat);
    void saveTextureAsRaw(int idx, const char* path);
    void saveTextureAsRaw(int idx, const wchar_t* path);
    void getTaskQCounters(unsigned int* taskPushCalled, unsigned int* taskPushSpinned, unsigned int* taskPopCalled,
                          unsigned int* taskPopSpinned);

    // access for callers that read results
    void finish(void);
    void readDisplay(void* bgraTopDown);
    psgpu_ctx* context() { return m_ctx; }

private:
    void check(int rc) const;
    void flushMirrors(void);                           // written getTexture mirrors -> GPU
    void hostPostProcess(PuresoftPostProcessor* processor);
    int fboTexture(PuresoftFBO* fbo);                  // the GPU texture standing for a host FBO
    void fboToGpu(PuresoftFBO* fbo, int tex);
    void fboFromGpu(PuresoftFBO* fbo, int tex);
    // bytes of texture idx, the depth buffer (-1) or the display (-2), level 0, layer 0, in memory row order
    void grab(int idx, std::vector<unsigned char>& bytes, int& w, int& scan, int& h, int& elem);
    psgpu_ctx* m_ctx = nullptr;
    int m_depthTex = -1;
    int m_width, m_height;
    std::unique_ptr<PuresoftHeadlessRenderer> m_ownRenderer;
    PuresoftHeadlessRenderer* m_renderer;
    std::vector<PuresoftProcessor*> m_processors;
    std::vector<std::vector<PuresoftVBO*>> m_vaoSlots;
    std::map<std::pair<int, int>, std::vector<unsigned char>> m_mirrors;   // (texture, layer) -> texels
    std::set<std::pair<int, int>> m_dirty;                               // handed out since the last upload
    std::map<PuresoftFBO*, int> m_fboTex;
    PuresoftFBO* m_fbo0 = nullptr;                                       // host FBO bound to slot 0, or none
    std::map<PuresoftPostProcessor*, std::unique_ptr<PsPostCall>> m_postCalls;
};
