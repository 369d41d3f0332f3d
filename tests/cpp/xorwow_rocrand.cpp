/*
 * Cross-check of the xorwow recurrence and jump-ahead against rocRAND's
 * open-source xorwow engine (rocrand/rocrand_xorwow.h, MIT licence): the state
 * is set directly (rocRAND's own seeding is not used), skipped with rocRAND's
 * discard(), and stepped with next().
 *   usage: xorwow_rocrand x0 x1 x2 x3 x4 d skip n
 *   prints the state after the skip, then n output words
 */
#include <rocrand/rocrand_xorwow.h>
#include <cstdio>
#include <cstdlib>

struct Probe : rocrand_device::xorwow_engine {
    Probe(const unsigned* x, unsigned d) : xorwow_engine(0, 0, 0)
    {
        for (int i = 0; i < 5; ++i) m_state.x[i] = x[i];
        m_state.d = d;
    }
    void dump() const
    {
        printf("%u %u %u %u %u %u\n", m_state.x[0], m_state.x[1], m_state.x[2], m_state.x[3], m_state.x[4], m_state.d);
    }
};

int main(int argc, char** argv)
{
    if (argc != 9) return 2;
    unsigned x[5];
    for (int i = 0; i < 5; ++i) x[i] = (unsigned)strtoul(argv[1 + i], nullptr, 10);
    const unsigned d = (unsigned)strtoul(argv[6], nullptr, 10);
    const unsigned long long skip = strtoull(argv[7], nullptr, 10);
    const int n = atoi(argv[8]);
    Probe p(x, d);
    p.discard(skip);
    p.dump();
    for (int i = 0; i < n; ++i) printf("%u\n", p.next());
    return 0;
}
