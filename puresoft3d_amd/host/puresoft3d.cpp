// puresoft3d.cpp -- PuresoftPipeline facade over libpsgpu (see puresoft3d.h).
// Object pools and argument checks follow pipeline.cpp, prog.cpp and tex.cpp;
// rendering goes through the C ABI.
#include "puresoft3d.h"

#include <string.h>

#include <new>
#include <string>

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

void PuresoftPipeline::destroyTexture(int idx) { check(psgpu_texture_destroy(m_ctx, idx)); }

void PuresoftPipeline::getTexture(int idx, PURESOFTIMGBUFF32* image, PuresoftFBO::LAYER layer)
{
    int w, s, h, e;
    check(psgpu_texture_info(m_ctx, idx, 0, &w, &s, &h, &e));
    if (image->pixels) check(psgpu_texture_download(m_ctx, idx, (int)layer, 0, image->pixels, (size_t)s * h));
    image->width = w;
    image->scanline = s;
    image->height = h;
    image->elemLen = e;
}

void PuresoftPipeline::setTexturePixels(int idx, const void* pixels, PuresoftFBO::LAYER layer)
{
    int w, s, h, e;
    check(psgpu_texture_info(m_ctx, idx, 0, &w, &s, &h, &e));
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
void PuresoftPipeline::setDepth(int idx) { check(psgpu_depth_set(m_ctx, idx)); }
void PuresoftPipeline::setFBO(int idx, int tex) { check(psgpu_target_set(m_ctx, idx, tex)); }

void PuresoftPipeline::setUniform(int idx, const void* data, size_t len)
{
    if (idx < 0 || idx >= 1024) throw std::out_of_range("PuresoftPipeline::setUniform");
    check(psgpu_uniform_set(m_ctx, idx, data, len));
}

void PuresoftPipeline::drawVAO(int vao, bool) { check(psgpu_draw(m_ctx, vao)); }
void PuresoftPipeline::enable(int b) { check(psgpu_behavior_enable(m_ctx, b)); }
void PuresoftPipeline::disable(int b) { check(psgpu_behavior_disable(m_ctx, b)); }
void PuresoftPipeline::clearDepth(float f) { check(psgpu_clear_depth(m_ctx, f)); }
void PuresoftPipeline::clearColour(PURESOFTBGRA c) { check(psgpu_clear_colour(m_ctx, c.i32)); }
void PuresoftPipeline::finish(void) { check(psgpu_sync(m_ctx)); }

void PuresoftPipeline::readDisplay(void* out) { check(psgpu_read_colour(m_ctx, out, (size_t)m_width * m_height * 4)); }

static int post_trampoline(void* user, const psgpu_post_frame* frame, void* stream)
{
    return static_cast<PuresoftPostProcessor*>(user)->process(*frame, stream);
}

void PuresoftPipeline::postProcess(PuresoftPostProcessor* processor)
{
    // post.cpp:3-18: after the draws so far, over the slot-0 target and the depth buffer
    if (!processor) throw std::invalid_argument("postProcess: no processor");
    check(psgpu_post_process(m_ctx, post_trampoline, processor));
}

void PuresoftPipeline::swapBuffers(void)
{
    // the display becomes visible through the renderer's buffer (pipeline.cpp:314-322)
    readDisplay(m_renderer->buffer());
}
