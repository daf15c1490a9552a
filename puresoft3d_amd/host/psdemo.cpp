// psdemo.cpp -- the reference demos' resource idioms through the facade,
// unchanged in shape:
//  * a 2D texture created from a picture loader's description (no pixels
//    yet), then filled through the pointer getTexture hands out
//    (test/scenobj.cpp:165-183, test2/loadscene.cpp:349-366);
//  * a six-layer skybox created from a NULL image, each layer filled through
//    getTexture's pointer for that layer (test/scenobj.cpp:185-219);
//  * a post-processor of the reference's form (test2/testpost.h,
//    test2/testpost.cpp:9-44): every 8 bytes of each of the thread's rows
//    plus 50 per byte, wrapping (paddb).
// It renders one frame (clear, the skybox, a textured DEF01 quad, the
// post-process, swapBuffers) and writes the presented BGRA8 frame.
//   psdemo OUT.bgra WIDTH HEIGHT [FBO.bgra]   (WIDTH even, as the post-processor's 2-pixel steps need)
// With FBO.bgra it then draws the quad again into a host FBO bound by
// setFBO(0, fbo) (pipeline.cpp:260-275), its memory first filled with 0x11
// bytes and the depth cleared, and writes the FBO's memory (bottom-up rows).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <wchar.h>

#include <map>
#include <string>

#include "puresoft3d.h"

// ---- a picture loader with the reference's interface (picldr.h) -----------------------
// Instead of decoding a file it makes a seeded picture: the file name picks
// the seed and the size.  Texel (row, col) of memory holds
//   B = 7 col + 3 row + s, G = 5 col + 11 row + 2 s, R = 13 col + row + 3 s (mod 256), A = 255.
class StubPictureLoader {
public:
    void loadFromFile(const wchar_t* path, PURESOFTIMGBUFF32* imageInfo)
    {
        m_seed = 0;
        for (const wchar_t* p = path; *p; p++) m_seed = (m_seed * 31 + (unsigned)*p) & 255;
        bool face = wcsncmp(path, L"sky", 3) == 0;
        m_w = face ? 32 : 64;
        m_h = 32;
        if (imageInfo) {   // picldr.cpp:49-56: the description, no pixels
            imageInfo->width = m_w;
            imageInfo->scanline = m_w * 4;
            imageInfo->height = m_h;
            imageInfo->elemLen = 4;
            imageInfo->pixels = NULL;
        }
    }
    void retrievePixel(PURESOFTIMGBUFF32* image)
    {
        unsigned char* row = (unsigned char*)image->pixels;
        for (unsigned y = 0; y < image->height; y++, row += image->scanline)
            for (unsigned x = 0; x < image->width; x++) {
                row[4 * x + 0] = (unsigned char)(7 * x + 3 * y + m_seed);
                row[4 * x + 1] = (unsigned char)(5 * x + 11 * y + 2 * m_seed);
                row[4 * x + 2] = (unsigned char)(13 * x + y + 3 * m_seed);
                row[4 * x + 3] = 255;
            }
    }
    void close(void) {}

private:
    unsigned m_seed = 0, m_w = 0, m_h = 0;
};

// ---- test2/testpost.h ---------------------------------------------------------------
class PP_DepthofField : public PuresoftPostProcessor {
public:
    void process(int threadIndex, int threadCount, PuresoftFBO* frame, PuresoftFBO* depth);
};

void PP_DepthofField::process(int threadIndex, int threadCount, PuresoftFBO* frame, PuresoftFBO* depth)
{
    uintptr_t frameBuffer = (uintptr_t)frame->getBuffer();
    int scanline = frame->getScanline();
    frameBuffer += threadIndex * scanline;
    for (int y = threadIndex; y < frame->getHeight(); y += threadCount) {
        PURESOFTBGRA* row = (PURESOFTBGRA*)frameBuffer;
        for (int x = 0; x < frame->getWidth(); x += 2) {
            unsigned char* b = (unsigned char*)row;
            for (int k = 0; k < 8; k++) b[k] = (unsigned char)(b[k] + 50);   // paddb
            row += 2;
        }
        frameBuffer += scanline * threadCount;
    }
}

// ---- test/scenobj.cpp:165-219, with the stub loader -------------------------------------
static std::map<std::string, int> g_textures;

static int findOrCreateTexture(PuresoftPipeline& pipeline, const char* pictureFile)
{
    auto it = g_textures.find(pictureFile);
    if (it != g_textures.end()) return it->second;
    std::wstring w(pictureFile, pictureFile + strlen(pictureFile));
    PURESOFTIMGBUFF32 image;
    StubPictureLoader picLoader;
    picLoader.loadFromFile(w.c_str(), &image);
    int tex = pipeline.createTexture(&image, 0, PuresoftFBO::WRAP);
    pipeline.getTexture(tex, &image);
    picLoader.retrievePixel(&image);
    picLoader.close();
    return g_textures[pictureFile] = tex;
}

static int findOrCreateSkybox(PuresoftPipeline& pipeline, const char* pictureFiles[6])
{
    std::string name;
    for (int i = 0; i < 6; i++) name += std::string(pictureFiles[i]) + ";";
    auto it = g_textures.find(name);
    if (it != g_textures.end()) return it->second;
    PURESOFTIMGBUFF32 image;
    StubPictureLoader picLoader;
    std::wstring w0(pictureFiles[0], pictureFiles[0] + strlen(pictureFiles[0]));
    picLoader.loadFromFile(w0.c_str(), &image);
    image.pixels = NULL;
    int tex = pipeline.createTexture(&image, 5);
    pipeline.getTexture(tex, &image, PuresoftFBO::LAYER_XPOS);
    picLoader.retrievePixel(&image);
    picLoader.close();
    for (int i = 1; i < 6; i++) {
        std::wstring w(pictureFiles[i], pictureFiles[i] + strlen(pictureFiles[i]));
        picLoader.loadFromFile(w.c_str(), &image);
        pipeline.getTexture(tex, &image, (PuresoftFBO::LAYER)i);
        picLoader.retrievePixel(&image);
        picLoader.close();
    }
    return g_textures[name] = tex;
}

// the view of the skybox: half a radian about y, column-major (the cosine and
// sine as float literals, so that other programs can rebuild the same bits)
static const float kView[16] = {0.87758256f, 0, -0.47942554f, 0, 0, 1, 0, 0, 0.47942554f, 0, 0.87758256f, 0, 0, 0, 0, 1};
static const char* kSkyFiles[6] = {"sky_rt.png", "sky_lf.png", "sky_up.png", "sky_dn.png", "sky_ft.png", "sky_bk.png"};

int main(int argc, char** argv)
{
    if (argc != 4 && argc != 5) {
        fprintf(stderr, "usage: psdemo OUT.bgra WIDTH HEIGHT [FBO.bgra]\n");
        return 2;
    }
    const int W = atoi(argv[2]), H = atoi(argv[3]);
    try {
        PuresoftPipeline pipeline(0, W, H);
        int sky = findOrCreateSkybox(pipeline, kSkyFiles);
        int quadTex = findOrCreateTexture(pipeline, "quad.png");
        if (findOrCreateTexture(pipeline, "quad.png") != quadTex) throw std::logic_error("texture cache");

        int vp4 = pipeline.addProcessor(new VertexProcesserDEF04), ip4 = pipeline.addProcessor(new InterpolationProcessorDEF04),
            fp4 = pipeline.addProcessor(new FragmentProcessorDEF04);
        int vp1 = pipeline.addProcessor(new VertexProcesserDEF01), ip1 = pipeline.addProcessor(new InterpolationProcessorDEF01),
            fp1 = pipeline.addProcessor(new FragmentProcessorDEF01);
        int skyProg = pipeline.createProgramme(vp4, ip4, fp4), quadProg = pipeline.createProgramme(vp1, ip1, fp1);

        // the skybox: a full-view quad at the far end (FullViewBuilder, test/testobjs.cpp:185)
        const float z = 0.999f;
        float skyPos[6][4] = {{-1, -1, z, 1}, {1, -1, z, 1}, {1, 1, z, 1}, {-1, -1, z, 1}, {1, 1, z, 1}, {-1, 1, z, 1}};
        int skyVao = pipeline.createVAO();
        pipeline.attachVBO(skyVao, 0, new PuresoftVBO(16, 6));
        pipeline.getVBO(skyVao, 0)->updateContent(skyPos);

        // a textured quad in front of it, lit from the upper left
        const float a = 0.6f;
        float pos[6][4] = {{-a, -a, 0.2f, 1}, {a, -a, 0.2f, 1}, {a, a, 0.2f, 1}, {-a, -a, 0.2f, 1}, {a, a, 0.2f, 1}, {-a, a, 0.2f, 1}};
        float nrm[6][4], uv[6][2];
        for (int i = 0; i < 6; i++) {
            nrm[i][0] = 0; nrm[i][1] = 0; nrm[i][2] = 1; nrm[i][3] = 0;
            uv[i][0] = (pos[i][0] + a) / (2 * a);
            uv[i][1] = (pos[i][1] + a) / (2 * a);
        }
        int quadVao = pipeline.createVAO();
        pipeline.attachVBO(quadVao, 0, new PuresoftVBO(16, 6));
        pipeline.attachVBO(quadVao, 3, new PuresoftVBO(16, 6));
        pipeline.attachVBO(quadVao, 4, new PuresoftVBO(8, 6));
        pipeline.getVBO(quadVao, 0)->updateContent(pos);
        pipeline.getVBO(quadVao, 3)->updateContent(nrm);
        pipeline.getVBO(quadVao, 4)->updateContent(uv);

        float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        float light[4] = {-3.0f, 4.0f, 6.0f, 0.0f}, camera[4] = {0.0f, 0.0f, 5.0f, 0.0f};
        PURESOFTBGRA bg;
        bg.i32 = 0x00203040;
        PP_DepthofField depthofField;

        pipeline.clearDepth();
        pipeline.clearColour(bg);
        pipeline.useProgramme(skyProg);
        pipeline.setUniform(0, I, sizeof(I));
        pipeline.setUniform(1, kView, sizeof(kView));
        pipeline.setUniform(2, &sky, sizeof(sky));
        pipeline.drawVAO(skyVao);
        pipeline.useProgramme(quadProg);
        pipeline.setUniform(3, I, sizeof(I));
        pipeline.setUniform(4, I, sizeof(I));
        pipeline.setUniform(5, I, sizeof(I));
        pipeline.setUniform(7, light, sizeof(light));
        pipeline.setUniform(8, camera, sizeof(camera));
        pipeline.setUniform(9, &quadTex, sizeof(quadTex));
        pipeline.drawVAO(quadVao);
        pipeline.postProcess(&depthofField);
        PuresoftHeadlessRenderer rndr(W, H);
        pipeline.setRenderer(&rndr);
        pipeline.swapBuffers();
        FILE* f = fopen(argv[1], "wb");
        if (!f) throw std::runtime_error("cannot write the output");
        fwrite(rndr.buffer(), 1, (size_t)W * H * 4, f);
        fclose(f);
        if (argc == 5) {
            PuresoftFBO target(W, W * 4, H, 4);
            memset(target.getBuffer(), 0x11, target.getBytes());
            pipeline.setFBO(0, &target);
            pipeline.clearDepth();
            pipeline.drawVAO(quadVao);
            pipeline.setFBO(0, NULL);
            FILE* g = fopen(argv[4], "wb");
            if (!g) throw std::runtime_error("cannot write the FBO output");
            fwrite(target.getBuffer(), 1, target.getBytes(), g);
            fclose(g);
        }
    } catch (const std::exception& e) {
        fprintf(stderr, "psdemo: %s\n", e.what());
        return 1;
    }
    return 0;
}
