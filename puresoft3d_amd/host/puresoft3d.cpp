// puresoft3d.cpp -- PuresoftPipeline facade over libpsgpu (see puresoft3d.h).
// Object pools and argument checks follow pipeline.cpp, prog.cpp and tex.cpp;
// rendering goes through the C ABI.
#include "puresoft3d.h"

#include <stdio.h>
#include <string.h>

#include <new>
#include <string>
#include <thread>

PuresoftFBO::PuresoftFBO(int width, int scanline, int height, int elemLen, bool topDown, void* externalBuffer,
                         WRAPMODE wrapMode, int extraLayers)
    : m_width(width), m_scanline(scanline), m_height(height), m_elemLen(elemLen), m_topDown(topDown),
      m_wrap(wrapMode), m_extraLayers(extraLayers), m_buffer(nullptr)
{
    // fbo.cpp:11-74
    if (width < 1 || height < 1 || scanline < width * elemLen || (elemLen != 1 && elemLen != 4))
        throw std::invalid_argument("PuresoftFBO::PuresoftFBO");
    setBuffer(externalBuffer);
}

void PuresoftFBO::setBuffer(void* externalBuffer)
{
    if (externalBuffer) {
        m_buffer = externalBuffer;
        return;
    }
    m_own.assign(getBytes(), 0);
    m_buffer = m_own.data();
}

PuresoftVBO::PuresoftVBO(size_t unitBytes, size_t unitCount)
    : m_unitBytes(unitBytes), m_unitCount(unitCount), m_host(unitBytes * unitCount)
{
    if (!unitBytes || !unitCount) throw std::invalid_argument("PuresoftVBO::PuresoftVBO");
}

void PuresoftVBO::updateContent(const void* src)
{
    memcpy(m_host.data(), src, m_host.size());   // vbo.cpp:28-31
    if (m_ctx && m_handle >= 0) {
        int rc = psgpu_vbo_update(m_ctx, m_handle, m_host.data());
        if (rc) throw std::runtime_error(psgpu_last_error(m_ctx));
    }
}

void PuresoftPipeline::check(int rc) const
{
    if (rc == PSGPU_OK) return;
    std::string msg = psgpu_last_error(m_ctx);
    switch (rc) {
    case PSGPU_ERR_RANGE: throw std::out_of_range(msg);
    case PSGPU_ERR_INVALID: throw std::invalid_argument(msg);
    case PSGPU_ERR_NOMEM: throw std::bad_alloc();
    default: throw std::runtime_error(msg);
    }
}

PuresoftPipeline::PuresoftPipeline(uintptr_t, int deviceWidth, int deviceHeight, PuresoftHeadlessRenderer* rndr, int device)
    : m_width(deviceWidth), m_height(deviceHeight), m_renderer(rndr)
{
    int rc = psgpu_create(device, deviceWidth, deviceHeight, 0, &m_ctx);
    if (rc) throw std::runtime_error("psgpu_create failed");
    if (!m_renderer) {
        m_ownRenderer.reset(new PuresoftHeadlessRenderer(deviceWidth, deviceHeight));
        m_renderer = m_ownRenderer.get();
    }
}

struct PsPostCall {
    PuresoftPostProcessor* proc;
    bool host = false;   // the processor has the reference's form only
};

PuresoftPipeline::~PuresoftPipeline()
{
    // the pipeline owns processors and attached VBOs (pipeline.cpp:81-87, 183-205)
    for (auto* p : m_processors) delete p;
    for (auto& slots : m_vaoSlots)
        for (auto* v : slots) delete v;
    psgpu_destroy(m_ctx);
}

// ---- textures (tex.cpp) ----------------------------------------------------------
int PuresoftPipeline::createTexture(const PURESOFTIMGBUFF32* image, int extraLayers, PuresoftFBO::WRAPMODE mode,
                                    int mipLevels)
{
    int idx = -1;
    check(psgpu_texture_create(m_ctx, (int)image->width, (int)image->scanline, (int)image->height, (int)image->elemLen,
                               extraLayers, (int)mode, mipLevels, image->pixels, &idx));
    return idx;
}

void PuresoftPipeline::destroyTexture(int idx)
{
    check(psgpu_texture_destroy(m_ctx, idx));
    for (int l = 0; l < PuresoftFBO::LAYER_MAX; l++) {
        m_mirrors.erase({idx, l});
        m_dirty.erase({idx, l});
    }
}

void PuresoftPipeline::getTexture(int idx, PURESOFTIMGBUFF32* image, PuresoftFBO::LAYER layer)
{
    int w, s, h, e;
    check(psgpu_texture_info(m_ctx, idx, 0, &w, &s, &h, &e));
    if ((int)layer < 0 || (int)layer >= PuresoftFBO::LAYER_MAX) throw std::out_of_range("PuresoftPipeline::getTexture");
    // the callers write texels through the pointer (test/scenobj.cpp:178-180,
    // 203-215; test2/loadscene.cpp:363): hand out a mirror holding the layer's
    // current texels, uploaded again before the GPU next uses textures
    std::vector<unsigned char>& m = m_mirrors[{idx, (int)layer}];
    m.resize((size_t)s * h);
    if (!m_dirty.count({idx, (int)layer})) check(psgpu_texture_download(m_ctx, idx, (int)layer, 0, m.data(), m.size()));
    m_dirty.insert({idx, (int)layer});
    image->width = w;
    image->scanline = s;
    image->height = h;
    image->elemLen = e;
    image->pixels = m.data();
}

void PuresoftPipeline::flushMirrors(void)
{
    std::set<int> mipped;
    for (const auto& k : m_dirty) {
        std::vector<unsigned char>& m = m_mirrors[k];
        check(psgpu_texture_upload(m_ctx, k.first, k.second, 0, m.data(), m.size()));
        mipped.insert(k.first);
    }
    m_dirty.clear();
    for (int idx : mipped) {   // a mip chain follows its level 0
        int w, s, h, e;
        if (psgpu_texture_info(m_ctx, idx, 1, &w, &s, &h, &e) == PSGPU_OK) check(psgpu_texture_generate_mips(m_ctx, idx));
    }
}

void PuresoftPipeline::setTexturePixels(int idx, const void* pixels, PuresoftFBO::LAYER layer)
{
    int w, s, h, e;
    check(psgpu_texture_info(m_ctx, idx, 0, &w, &s, &h, &e));
    auto it = m_mirrors.find({idx, (int)layer});
    if (it != m_mirrors.end()) {   // a handed-out mirror follows the texture
        memcpy(it->second.data(), pixels, it->second.size());
        m_dirty.erase({idx, (int)layer});
    }
    check(psgpu_texture_upload(m_ctx, idx, (int)layer, 0, pixels, (size_t)s * h));
}

void PuresoftPipeline::generateMipmaps(int idx) { check(psgpu_texture_generate_mips(m_ctx, idx)); }

// ---- processors and programmes (prog.cpp) -----------------------------------------
int PuresoftPipeline::addProcessor(PuresoftProcessor* proc)
{
    size_t vacant = 0;
    for (; vacant < m_processors.size(); vacant++)
        if (!m_processors[vacant]) break;
    if (vacant == m_processors.size()) m_processors.push_back(nullptr);
    m_processors[vacant] = proc;
    return (int)vacant;
}

void PuresoftPipeline::destroyProcessor(int idx)
{
    if (idx < 0 || idx >= (int)m_processors.size()) throw std::out_of_range("PuresoftPipeline::destroyProcessor");
    delete m_processors[idx];
    m_processors[idx] = nullptr;
}

int PuresoftPipeline::createProgramme(int vid, int iid, int fid)
{
    auto kind = [&](int i, const char* what) {
        if (i < 0 || i >= (int)m_processors.size() || !m_processors[i])
            throw std::out_of_range(std::string("PuresoftPipeline::createProgramme, ") + what);
        return m_processors[i]->gpuKind();
    };
    int vk = kind(vid, "vid"), ik = kind(iid, "iid"), fk = kind(fid, "fid");
    if (!dynamic_cast<PuresoftVertexProcessor*>(m_processors[vid]) ||
        !dynamic_cast<PuresoftInterpolationProcessor*>(m_processors[iid]) ||
        !dynamic_cast<PuresoftFragmentProcessor*>(m_processors[fid]))
        throw std::invalid_argument("PuresoftPipeline::createProgramme: processor roles");
    int prog = -1;
    check(psgpu_program_create(m_ctx, vk, ik, fk, &prog));
    return prog;
}

void PuresoftPipeline::destroyProgramme(int idx) { check(psgpu_program_destroy(m_ctx, idx)); }
void PuresoftPipeline::useProgramme(int idx) { check(psgpu_program_use(m_ctx, idx)); }

// ---- vertex streams (pipeline.cpp:118-205) ------------------------------------------
int PuresoftPipeline::createVAO(void)
{
    int vao = -1;
    check(psgpu_vao_create(m_ctx, &vao));
    if (vao >= (int)m_vaoSlots.size()) m_vaoSlots.resize(vao + 1);
    m_vaoSlots[vao].assign(16, nullptr);
    return vao;
}

PuresoftVBO* PuresoftPipeline::attachVBO(int vao, int idx, PuresoftVBO* vbo)
{
    if (vao < 0 || vao >= (int)m_vaoSlots.size() || m_vaoSlots[vao].empty()) throw std::out_of_range("PuresoftPipeline::attachVBO vao");
    if (idx < 0 || idx >= 16) throw std::out_of_range("PuresoftPipeline::attachVBO idx");
    if (vbo && vbo->m_handle < 0) {
        check(psgpu_vbo_create(m_ctx, vbo->m_unitBytes, vbo->m_unitCount, &vbo->m_handle));
        vbo->m_ctx = m_ctx;
        check(psgpu_vbo_update(m_ctx, vbo->m_handle, vbo->m_host.data()));
    }
    int prev = -1;
    check(psgpu_vao_attach(m_ctx, vao, idx, vbo ? vbo->m_handle : -1, &prev));
    PuresoftVBO* old = m_vaoSlots[vao][idx];
    m_vaoSlots[vao][idx] = vbo;
    return old;
}

PuresoftVBO* PuresoftPipeline::detachVBO(int vao, int idx) { return attachVBO(vao, idx, nullptr); }

PuresoftVBO* PuresoftPipeline::getVBO(int vao, int idx)
{
    if (vao < 0 || vao >= (int)m_vaoSlots.size() || m_vaoSlots[vao].empty()) throw std::out_of_range("PuresoftPipeline::getVBO vao");
    if (idx < 0 || idx >= 16) throw std::out_of_range("PuresoftPipeline::getVBO idx");
    return m_vaoSlots[vao][idx];
}

void PuresoftPipeline::destroyVAO(int vao)
{
    if (vao < 0 || vao >= (int)m_vaoSlots.size()) throw std::out_of_range("PuresoftPipeline::destroyVAO");
    check(psgpu_vao_destroy(m_ctx, vao));
    for (auto* v : m_vaoSlots[vao]) delete v;
    m_vaoSlots[vao].clear();
}

// ---- rendering ------------------------------------------------------------------------
void PuresoftPipeline::setViewport(int width, int height, uintptr_t) { check(psgpu_viewport(m_ctx, width, height)); }
void PuresoftPipeline::setDepth(int idx)
{
    check(psgpu_depth_set(m_ctx, idx));
    m_depthTex = idx;
}

void PuresoftPipeline::setRenderer(PuresoftHeadlessRenderer* rndr)
{
    // pipeline.cpp:241-258: a null renderer makes the pipeline create its own
    if (rndr) {
        m_renderer = rndr;
    } else {
        m_ownRenderer.reset(new PuresoftHeadlessRenderer(m_width, m_height));
        m_renderer = m_ownRenderer.get();
    }
}

// ---- debug api (dbg.cpp:3-61) ---------------------------------------------------
void PuresoftPipeline::grab(int idx, std::vector<unsigned char>& bytes, int& w, int& scan, int& h, int& elem)
{
    flushMirrors();
    if (idx == -2) {   // the display, rows top-down as it lies in memory
        w = m_width; h = m_height; elem = 4; scan = w * 4;
        bytes.resize((size_t)scan * h);
        check(psgpu_read_colour(m_ctx, bytes.data(), bytes.size()));
        return;
    }
    if (idx == -1 && m_depthTex < 0) {   // the default depth buffer, FP32
        w = m_width; h = m_height; elem = 4; scan = (int)(w / 4.0f + 0.5f) * 4 * 4;
        bytes.resize((size_t)scan * h);
        check(psgpu_read_depth(m_ctx, bytes.data(), bytes.size()));
        return;
    }
    int t = idx == -1 ? m_depthTex : idx;
    if (t < 0) throw std::out_of_range("PuresoftPipeline::saveTexture");
    check(psgpu_texture_info(m_ctx, t, 0, &w, &scan, &h, &elem));
    bytes.resize((size_t)scan * h);
    check(psgpu_texture_download(m_ctx, t, 0, 0, bytes.data(), bytes.size()));
}

static std::string narrow(const wchar_t* p)
{
    std::string s;
    for (; p && *p; p++) s.push_back((char)*p);
    return s;
}

void PuresoftPipeline::saveTextureAsRaw(int idx, const char* path)
{
    // PuresoftFBO::saveAsRawFile (fbo.cpp:542-550): the buffer's bytes as they are
    std::vector<unsigned char> b;
    int w, scan, h, e;
    grab(idx, b, w, scan, h, e);
    FILE* f = fopen(path, "wb");
    if (!f) return;   // the reference ignores an unopenable path
    fwrite(b.data(), 1, b.size(), f);
    fclose(f);
}

void PuresoftPipeline::saveTextureAsRaw(int idx, const wchar_t* path) { saveTextureAsRaw(idx, narrow(path).c_str()); }

static void put16(FILE* f, unsigned v) { fputc(v & 255, f); fputc((v >> 8) & 255, f); }
static void put32(FILE* f, unsigned v) { put16(f, v & 0xffff); put16(f, v >> 16); }

void PuresoftPipeline::saveTexture(int idx, const char* path, bool dataIsFloat)
{
    // PuresoftFBO::saveAsBmpFile (fbo.cpp:473-540), written as a well-formed
    // BMP: byte texels as 32-bit BGRA with memory row h-1 first, FP32 texels
    // as 8-bit grey of value * 255 under a grey palette, rows in memory order.
    std::vector<unsigned char> b;
    int w, scan, h, e;
    grab(idx, b, w, scan, h, e);
    FILE* f = fopen(path, "wb");
    if (!f) return;
    if (!dataIsFloat) {
        unsigned rowb = (unsigned)w * e, img = rowb * h;
        fputc('B', f); fputc('M', f); put32(f, 54 + img); put32(f, 0); put32(f, 54);
        put32(f, 40); put32(f, w); put32(f, h); put16(f, 1); put16(f, e * 8); put32(f, 0); put32(f, img);
        put32(f, 2835); put32(f, 2835); put32(f, 0); put32(f, 0);
        for (int y = h - 1; y >= 0; y--) fwrite(b.data() + (size_t)y * scan, 1, rowb, f);
    } else {
        int chans = e / 4, line = (w * chans + 3) / 4 * 4;
        unsigned img = (unsigned)line * h, off = 54 + 1024;
        fputc('B', f); fputc('M', f); put32(f, off + img); put32(f, 0); put32(f, off);
        put32(f, 40); put32(f, w); put32(f, h); put16(f, 1); put16(f, chans * 8); put32(f, 0); put32(f, img);
        put32(f, 2835); put32(f, 2835); put32(f, 256); put32(f, 0);
        for (int i = 0; i < 256; i++) { fputc(i, f); fputc(i, f); fputc(i, f); fputc(0, f); }
        std::vector<unsigned char> row(line, 0);
        for (int y = 0; y < h; y++) {
            const float* src = (const float*)(b.data() + (size_t)y * scan);
            for (int x = 0; x < w * chans; x++) row[x] = (unsigned char)(int)(src[x] * 255.0f);
            fwrite(row.data(), 1, line, f);
        }
    }
    fclose(f);
}

void PuresoftPipeline::saveTexture(int idx, const wchar_t* path, bool dataIsFloat)
{
    saveTexture(idx, narrow(path).c_str(), dataIsFloat);
}

void PuresoftPipeline::getTaskQCounters(unsigned int* taskPushCalled, unsigned int* taskPushSpinned,
                                        unsigned int* taskPopCalled, unsigned int* taskPopSpinned)
{
    // the reference reports zeros unless built with PROFILING (dbg.cpp:45-61);
    // there are no row-task queues here
    *taskPushCalled = *taskPushSpinned = *taskPopCalled = *taskPopSpinned = 0;
}
void PuresoftPipeline::setFBOTexture(int idx, int tex)
{
    check(psgpu_target_set(m_ctx, idx, tex));
    if (idx == 0) m_fbo0 = nullptr;
}

// the GPU texture standing for a host FBO (created the first time it is bound)
int PuresoftPipeline::fboTexture(PuresoftFBO* fbo)
{
    auto it = m_fboTex.find(fbo);
    if (it != m_fboTex.end()) return it->second;
    if (fbo->getElemLen() != 4) throw std::invalid_argument("PuresoftPipeline::setFBO: targets hold 4-byte texels");
    int tex = -1;
    check(psgpu_texture_create(m_ctx, fbo->getWidth(), fbo->getScanline(), fbo->getHeight(), 4, 0,
                               (int)fbo->getWrapMode(), 1, nullptr, &tex));
    m_fboTex[fbo] = tex;
    return tex;
}

// texture rows are bottom-up: a top-down FBO's rows go across mirrored
void PuresoftPipeline::fboToGpu(PuresoftFBO* fbo, int tex)
{
    const size_t scan = fbo->getScanline(), h = fbo->getHeight();
    const unsigned char* src = (const unsigned char*)fbo->getBuffer();
    if (!fbo->isTopDown()) {
        check(psgpu_texture_upload(m_ctx, tex, 0, 0, src, scan * h));
        return;
    }
    std::vector<unsigned char> rows(scan * h);
    for (size_t y = 0; y < h; y++) memcpy(&rows[y * scan], src + (h - 1 - y) * scan, scan);
    check(psgpu_texture_upload(m_ctx, tex, 0, 0, rows.data(), rows.size()));
}

void PuresoftPipeline::fboFromGpu(PuresoftFBO* fbo, int tex)
{
    const size_t scan = fbo->getScanline(), h = fbo->getHeight();
    unsigned char* dst = (unsigned char*)fbo->getBuffer();
    if (!fbo->isTopDown()) {
        check(psgpu_texture_download(m_ctx, tex, 0, 0, dst, scan * h));
        return;
    }
    std::vector<unsigned char> rows(scan * h);
    check(psgpu_texture_download(m_ctx, tex, 0, 0, rows.data(), rows.size()));
    for (size_t y = 0; y < h; y++) memcpy(dst + (h - 1 - y) * scan, &rows[y * scan], scan);
}

void PuresoftPipeline::setFBO(int idx, PuresoftFBO* fbo)
{
    if (idx < 0 || idx >= 16) throw std::out_of_range("PuresoftPipeline::setFBO");
    if (!fbo) {   // the display for slot 0, none for the others (pipeline.cpp:266-274)
        setFBOTexture(idx, PSGPU_DISPLAY);
        return;
    }
    int tex = fboTexture(fbo);
    check(psgpu_target_set(m_ctx, idx, tex));
    if (idx == 0) m_fbo0 = fbo;
}

void PuresoftPipeline::setUniform(int idx, const void* data, size_t len)
{
    if (idx < 0 || idx >= 1024) throw std::out_of_range("PuresoftPipeline::setUniform");
    check(psgpu_uniform_set(m_ctx, idx, data, len));
}

void PuresoftPipeline::drawVAO(int vao, bool)
{
    flushMirrors();
    if (!m_fbo0) {
        check(psgpu_draw(m_ctx, vao));
        return;
    }
    // a host FBO target: the draw sees and leaves its memory, as the
    // reference's synchronous drawVAO does (drawvao.cpp:122-132)
    int tex = m_fboTex.at(m_fbo0);
    fboToGpu(m_fbo0, tex);
    check(psgpu_draw(m_ctx, vao));
    fboFromGpu(m_fbo0, tex);
}
void PuresoftPipeline::enable(int b) { check(psgpu_behavior_enable(m_ctx, b)); }
void PuresoftPipeline::disable(int b) { check(psgpu_behavior_disable(m_ctx, b)); }
void PuresoftPipeline::clearDepth(float f) { check(psgpu_clear_depth(m_ctx, f)); }
void PuresoftPipeline::clearColour(PURESOFTBGRA c) { check(psgpu_clear_colour(m_ctx, c.i32)); }
void PuresoftPipeline::finish(void)
{
    flushMirrors();
    check(psgpu_sync(m_ctx));
}

void PuresoftPipeline::readDisplay(void* out)
{
    flushMirrors();
    check(psgpu_read_colour(m_ctx, out, (size_t)m_width * m_height * 4));
}

static int post_trampoline(void* user, const psgpu_post_frame* frame, void* stream)
{
    PsPostCall* pc = static_cast<PsPostCall*>(user);
    int rc = pc->proc->process(*frame, stream);
    if (rc == PuresoftPostProcessor::kHostForm) {   // the reference's form: run by the pipeline afterwards
        pc->host = true;
        return PSGPU_OK;
    }
    return rc;
}

void PuresoftPipeline::postProcess(PuresoftPostProcessor* processor)
{
    // post.cpp:3-18: after the draws so far, over the slot-0 target and the depth buffer
    if (!processor) throw std::invalid_argument("postProcess: no processor");
    finish();   // the draws before it are done: the callback below runs within the call
    std::unique_ptr<PsPostCall>& pc = m_postCalls[processor];
    if (!pc) pc.reset(new PsPostCall{processor});
    pc->host = false;
    check(psgpu_post_process(m_ctx, post_trampoline, pc.get()));
    if (pc->host) hostPostProcess(processor);
}

// the reference's form: the frame and the depth buffer as host FBOs, the
// processor called from threadCount threads (the reference's workers plus
// its main thread, post.cpp:3-18), both buffers written back
void PuresoftPipeline::hostPostProcess(PuresoftPostProcessor* processor)
{
    check(psgpu_sync(m_ctx));
    std::unique_ptr<PuresoftFBO> display, depth;
    PuresoftFBO* frame = m_fbo0;   // a host FBO target already holds the frame (drawVAO)
    if (!frame) {
        display.reset(new PuresoftFBO(m_width, m_width * 4, m_height, 4, true));   // the display, top-down
        frame = display.get();
        check(psgpu_read_colour(m_ctx, frame->getBuffer(), frame->getBytes()));
    }
    int dscan = (int)(m_width / 4.0f + 0.5f) * 4 * 4;   // pipeline.cpp:31
    if (m_depthTex < 0) {
        depth.reset(new PuresoftFBO(m_width, dscan, m_height, 4));
        check(psgpu_read_depth(m_ctx, depth->getBuffer(), depth->getBytes()));
    } else {
        int w, s, h, e;
        check(psgpu_texture_info(m_ctx, m_depthTex, 0, &w, &s, &h, &e));
        depth.reset(new PuresoftFBO(w, s, h, e));
        check(psgpu_texture_download(m_ctx, m_depthTex, 0, 0, depth->getBuffer(), depth->getBytes()));
    }
    unsigned hw = std::thread::hardware_concurrency();
    int n = (int)(hw < 2 ? 2 : (hw > 16 ? 16 : hw));
    std::vector<std::thread> workers;
    for (int i = 0; i < n - 1; i++)
        workers.emplace_back([=, &depth] { processor->process(i, n, frame, depth.get()); });
    processor->process(n - 1, n, frame, depth.get());
    for (auto& t : workers) t.join();
    if (m_fbo0) fboToGpu(frame, m_fboTex.at(m_fbo0));
    else check(psgpu_write_colour(m_ctx, frame->getBuffer(), frame->getBytes()));
    if (m_depthTex < 0) check(psgpu_write_depth(m_ctx, depth->getBuffer(), depth->getBytes()));
    else check(psgpu_texture_upload(m_ctx, m_depthTex, 0, 0, depth->getBuffer(), depth->getBytes()));
}

void PuresoftPipeline::swapBuffers(void)
{
    // the display becomes visible through the renderer's buffer (pipeline.cpp:314-322)
    readDisplay(m_renderer->buffer());
}
