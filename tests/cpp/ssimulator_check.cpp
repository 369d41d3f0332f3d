/*
 * Caller of the MI355X simulate path through the reference's simulator
 * protocol: compiled against the reference's own sysrap/SSimulator.h (the
 * compiler checks that SMI355XSimulator overrides every pure virtual with the
 * reference's signatures), then driven only through an SSimulator pointer the
 * way G4CXOpticks drives CSGOptiX (g4cx/G4CXOpticks.cc:470-511).
 *
 *   ssimulator_check <csgfoundry dir> <gensteps.npy> <seed> <out dir>
 *   writes photon.bin, seq.bin, hit.bin into <out dir> and prints
 *   "<n_photon> <n_hit> <timed> <desc>" then "after reset <n_gs> <n_photon> <n_hit>"
 *   and "render <value>".
 */
#include <cstdio>
#include <cstdlib>
#include <type_traits>

#include "SMI355XSimulator.hh"

static_assert(std::is_base_of<SSimulator, SMI355XSimulator>::value, "not an SSimulator");
static_assert(!std::is_abstract<SMI355XSimulator>::value, "a protocol method is not implemented");

static void dump(const std::string& path, const void* p, size_t bytes)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) std::exit(3);
    if (bytes) std::fwrite(p, 1, bytes, f);
    std::fclose(f);
}

int main(int argc, char** argv)
{
    if (argc != 5) return 2;
    mi_config cfg;
    mi_config_default(&cfg);
    cfg.seed = std::strtoull(argv[3], nullptr, 10);
    SMI355XSimulator* impl = SMI355XSimulator::Load(argv[1], &cfg);
    SSimulator* s = impl;                                   /* from here on, the protocol only */
    const std::vector<float> gs = mi_npy::load(argv[2]).as_f32();
    impl->sim().add_gensteps(gs.data(), (int64_t)gs.size() / 24);
    const double ms = s->simulate(0, false);
    const MI355XSimulator& m = impl->sim();
    const std::string out = argv[4];
    dump(out + "/photon.bin", m.photon(), (size_t)m.num_photon() * 64);
    dump(out + "/seq.bin", m.seq(), (size_t)m.num_photon() * 32);
    dump(out + "/hit.bin", m.hit(), (size_t)m.num_hit() * 64);
    std::printf("%lld %lld %d %s\n", (long long)m.num_photon(), (long long)m.num_hit(), ms > 0.0 ? 1 : 0, s->desc());
    s->reset(0);
    std::printf("after reset %lld %lld %lld\n", (long long)impl->sim().num_genstep(), (long long)m.num_photon(),
                (long long)m.num_hit());
    std::printf("render %g %g\n", s->render(), s->render_launch());
    delete s;                                               /* virtual destructor of the protocol */
    return 0;
}
