// Quantized wide-node check: every decoded plane p + q*step must be exact in
// FP32 (one fma) and every quantized child box must contain its FP32 box;
// empty slots must decode inverted.  Random boxes over 40 binary orders of
// magnitude of size and offsets up to 2^24, including flat (zero-extent) axes.
#include "bvh.hpp"
#include <cmath>
#include <cstdio>
#include <random>
using namespace eleven;
template <int W> int run(std::mt19937& g) {
    std::uniform_real_distribution<float> U(-1, 1);
    int bad = 0;
    for (int it = 0; it < 200000; it++) {
        WideNode<W> n;
        float scale = std::ldexp(1.0f, (int)(g() % 40) - 20), off = U(g) * std::ldexp(1.0f, (int)(g() % 30) - 5);
        int nc = 1 + g() % W;
        for (int i = 0; i < W; i++) {
            if (i >= nc) { n.child[i] = -1; n.lox[i]=n.loy[i]=n.loz[i]=INFINITY; n.hix[i]=n.hiy[i]=n.hiz[i]=-INFINITY; continue; }
            n.child[i] = i;
            float* lo[3] = {&n.lox[i], &n.loy[i], &n.loz[i]}; float* hi[3] = {&n.hix[i], &n.hiy[i], &n.hiz[i]};
            for (int a = 0; a < 3; a++) {
                float c = off + U(g) * scale, e = (g() % 5 == 0) ? 0.f : std::fabs(U(g)) * scale * ((g()%3==0)?1e-6f:1.f);
                *lo[a] = c - e; *hi[a] = c + e;
            }
        }
        QNode<W> q;
        if (!quantizeNode<W>(n, q)) { bad++; continue; }
        const float* lo[3] = {n.lox, n.loy, n.loz}; const float* hi[3] = {n.hix, n.hiy, n.hiz};
        for (int a = 0; a < 3; a++) {
            float s = std::ldexp(1.0f, (int)((q.exps >> (8 * a)) & 0xff) - 127);
            for (int i = 0; i < W; i++) {
                uint32_t ql = q.plane(a*2, i), qh = q.plane(a*2+1, i);
                float bl = std::fma((float)ql, s, q.p[a]), bh = std::fma((float)qh, s, q.p[a]);
                double el = (double)q.p[a] + (double)ql * s, eh = (double)q.p[a] + (double)qh * s;
                if ((double)bl != el || (double)bh != eh) { bad++; if (bad < 5) printf("inexact\n"); }
                if (n.child[i] == -1) { if (!(ql == 255 && qh == 0)) bad++; continue; }
                if (!(bl <= lo[a][i] && bh >= hi[a][i])) { bad++; if (bad < 5) printf("not containing a=%d %g %g vs %g %g\n", a, bl, bh, lo[a][i], hi[a][i]); }
            }
        }
    }
    return bad;
}
int main() { std::mt19937 g(1); int b4 = run<4>(g), b8 = run<8>(g); printf("bad4=%d bad8=%d\n", b4, b8); return b4 || b8; }
