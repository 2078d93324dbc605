// rayna_engine/tests/simple_renders.rs (sphere_colours) and tests/common/mod.rs (render_simple), through the
// C++ Renderer API of include/rayna/engine.hpp.  Colours come from a seeded generator (the engine's test draws
// them from thread_rng); the checks are the engine's: inside the sphere all black, just outside it all colour,
// facing away all white (exact fractions), and at z = -2.7 mostly colour with 20-40% white.
// Exit status 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <random>

#include "rayna/engine.hpp"

using namespace rayna;

namespace common {  // tests/common/mod.rs
const RenderOpts SIMPLE_RENDER_OPTIONS = {320, 320, 10, RenderMode::PBR, 5, 1, 0};
const std::size_t RENDERER_THREAD_COUNT = 4;

Image render_simple(StandardScene scene, Camera camera) {
    Renderer rend = Renderer::new_from(std::move(scene), camera, SIMPLE_RENDER_OPTIONS, RENDERER_THREAD_COUNT);
    return rend.render().img;
}
}  // namespace common

static int failures = 0;
#define CHECK(cond)                                                                          \
    do {                                                                                     \
        if (!(cond)) {                                                                       \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);                    \
            ++failures;                                                                      \
        }                                                                                    \
    } while (0)

static void sphere_colours_internal(ColourRgb target_col, Channel thresh) {
    const StandardScene scene{{SimpleObject::new_uncorrected(SphereMesh{Point3::ZERO(), 1.0},
                                                             LambertianMaterial{target_col})},
                              SkyboxInstance::WhiteSkybox};
    const Camera camera{Point3::ZERO(), Vector3{0., 0., 1.}, Angle::from_degrees(45.), 1., Angle::from_degrees(0.)};

    auto colours_eq = [](ColourRgb px, ColourRgb target, Channel th) {
        return std::fabs(target.r - px.r) < th && std::fabs(target.g - px.g) < th && std::fabs(target.b - px.b) < th;
    };
    struct Counts { Number colour, white, black, other; };
    auto count_colours = [&](const Image &img) {
        std::size_t colour = 0, white = 0, black = 0, other = 0, total = 0;
        for (const ColourRgb &px : img) {
            if (colours_eq(px, ColourRgb::WHITE(), thresh)) ++white;
            else if (colours_eq(px, ColourRgb::BLACK(), thresh)) ++black;
            else if (colours_eq(px, target_col, thresh)) ++colour;
            else ++other;
            ++total;
        }
        const Number n = (Number)total;
        return Counts{colour / n, white / n, black / n, other / n};
    };
    auto at = [&](Number z) {
        Camera c = camera;
        c.pos = Point3{0., 0., z};
        return c;
    };

    std::printf("current colour is: (%g, %g, %g)\n", target_col.r, target_col.g, target_col.b);
    std::printf("rendering inside sphere, expecting all black...\n");
    Counts c = count_colours(common::render_simple(scene, at(0.)));
    CHECK(c.colour == 0. && c.white == 0. && c.black == 1. && c.other == 0.);

    std::printf("rendering just next to sphere, expecting all colour...\n");
    c = count_colours(common::render_simple(scene, at(-1.01)));
    CHECK(c.colour == 1. && c.white == 0. && c.black == 0. && c.other == 0.);

    std::printf("rendering the sky, expecting all white...\n");
    c = count_colours(common::render_simple(scene, at(1.01)));
    CHECK(c.colour == 0. && c.white == 1. && c.black == 0. && c.other == 0.);

    std::printf("rendering near sphere, expecting mostly colour...\n");
    c = count_colours(common::render_simple(scene, at(-2.7)));
    CHECK(c.colour >= 0.6);
    CHECK(0.2 <= c.white && c.white <= 0.4);
    CHECK(c.other <= 0.05);
    CHECK(c.black == 0.);
}

int main() {
    const Channel THRESHOLD = 0.1f;
    const int COUNT = 10;
    std::mt19937_64 gen(20240607);
    std::uniform_real_distribution<float> u(0.01f + THRESHOLD, 0.99f - THRESHOLD);
    for (int i = 0; i < COUNT; ++i) sphere_colours_internal(ColourRgb{u(gen), u(gen), u(gen)}, THRESHOLD);

    // RendererCreateError for options the GPU renderer does not take
    bool threw = false;
    try {
        RenderOpts o = common::SIMPLE_RENDER_OPTIONS;
        o.ray_branching = 2;
        (void)Renderer::new_from(StandardScene{}, Camera{}, o, 1);
    } catch (const RendererCreateError &e) {
        threw = e.status == RAYNA_UNSUPPORTED;
    }
    CHECK(threw);
    std::printf(failures ? "sphere_colours: %d check(s) failed\n" : "sphere_colours: ok\n", failures);
    return failures ? 1 : 0;
}
