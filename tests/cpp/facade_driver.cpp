// Test driver of include/kissfft.hh: reads complex (or real) input from a file, applies
// one operation of the class, writes the result.
//   facade_driver c2c    N inverse stride in out     transform(in, out, 0, 1, stride)
//   facade_driver real   N inverse        in out     transform_real (2N reals in)
//   facade_driver many   N inverse howmany in out    transform_many (dense)
//   facade_driver assign N inverse        in out     construct (7, !inverse), copy, move, assign(N, inverse), transform
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "kissfft.hh"

typedef kissfft<float>::cpx_t cpx;

static std::vector<char> slurp(const char *p)
{
    std::vector<char> b;
    FILE *f = std::fopen(p, "rb");
    if (!f)
        std::exit(2);
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0)
        b.insert(b.end(), buf, buf + n);
    std::fclose(f);
    return b;
}

static void spill(const char *p, const void *d, size_t n)
{
    FILE *f = std::fopen(p, "wb");
    if (!f || std::fwrite(d, 1, n, f) != n)
        std::exit(3);
    std::fclose(f);
}

int main(int argc, char **argv)
{
    if (argc < 6)
        return 1;
    const std::string mode = argv[1];
    const size_t N = std::strtoul(argv[2], nullptr, 10);
    const bool inv = std::atoi(argv[3]) != 0;
    try {
        if (mode == "c2c") {
            const size_t stride = std::strtoul(argv[4], nullptr, 10);
            std::vector<char> in = slurp(argv[5]);
            std::vector<cpx> out(N);
            kissfft<float> f(N, inv);
            f.transform(reinterpret_cast<const cpx *>(in.data()), out.data(), 0, 1, stride);
            spill(argv[6], out.data(), N * sizeof(cpx));
        } else if (mode == "real") {
            std::vector<char> in = slurp(argv[4]);
            std::vector<cpx> out(N);
            kissfft<float> f(N, inv);
            f.transform_real(reinterpret_cast<const float *>(in.data()), out.data());
            spill(argv[5], out.data(), N * sizeof(cpx));
        } else if (mode == "many") {
            const size_t howmany = std::strtoul(argv[4], nullptr, 10);
            std::vector<char> in = slurp(argv[5]);
            std::vector<cpx> out(N * howmany);
            kissfft<float> f(N, inv);
            f.transform_many(reinterpret_cast<const cpx *>(in.data()), out.data(), howmany, N, N);
            spill(argv[6], out.data(), out.size() * sizeof(cpx));
        } else if (mode == "assign") {
            std::vector<char> in = slurp(argv[4]);
            std::vector<cpx> out(N);
            kissfft<float> a(7, !inv);
            kissfft<float> b(a);             // copy
            kissfft<float> c(std::move(b));  // move
            c.assign(N, inv);
            a = c;                           // copy assignment
            a.transform(reinterpret_cast<const cpx *>(in.data()), out.data());
            spill(argv[5], out.data(), N * sizeof(cpx));
        } else {
            return 1;
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 4;
    }
    return 0;
}
