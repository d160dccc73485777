// Leaf block preconditioner: octree partition and dense leaf inverses.
#include "precond.hpp"

#include <cmath>

#pragma clang fp contract(off)

namespace tdbem {

namespace {

struct Cube {
    double c[3];
    double len;
};

// Point-in-cube test with the reference's 1.0001 slack (DBEM1/geo.cpp:8-24).
bool Inside(const double* p, const Cube& cb)
{
    double l = 0.5 * cb.len;
    for (int i = 0; i < 3; ++i) {
        double d = cb.c[i] - p[i];
        if (d * d > l * l * 1.0001)
            return false;
    }
    return true;
}

// Child cube by octant bits x=1, y=2, z=4 (DBEM1/geo.cpp:27-40).
Cube Child(const Cube& fa, int pos)
{
    Cube ch;
    ch.len = 0.5 * fa.len;
    double half = 0.5 * ch.len;
    for (int i = 0, bit = 1; i < 3; ++i, bit <<= 1)
        ch.c[i] = fa.c[i] + ((bit & pos) == 0 ? -1 : 1) * half;
    return ch;
}

// Recursive split.  The reference keeps each cell's points in a linked list
// to which points are prepended, so a child's list is the reverse of the
// order in which its points leave the parent list.
void Split(const double* pts, const Cube& cb, std::vector<long>& list, long begin, long maxLeaf,
           LeafPreconditioner& out)
{
    long cnt = (long)list.size();
    if (cnt <= maxLeaf) {
        out.leafSize.push_back(cnt);
        out.leafBegin.push_back(begin);
        for (long i = 0; i < cnt; ++i)
            out.repid[(size_t)(begin + i)] = list[(size_t)i];
        return;
    }
    Cube ch[8];
    std::vector<long> sub[8];
    for (int i = 0; i < 8; ++i)
        ch[i] = Child(cb, i);
    for (long p : list) {
        const double* x = &pts[3 * p];
        int k = -1;
        for (int i = 0; i < 8; ++i)
            if (Inside(x, ch[i])) {
                k = i;
                break;
            }
        if (k < 0) {
            double best = Dist3(x, ch[0].c);
            k = 0;
            for (int i = 1; i < 8; ++i) {
                double d = Dist3(x, ch[i].c);
                if (best > d) {
                    best = d;
                    k = i;
                }
            }
        }
        sub[k].push_back(p);
    }
    long b = begin;
    for (int i = 0; i < 8; ++i) {
        if (sub[i].empty())
            continue;
        std::vector<long> rev(sub[i].rbegin(), sub[i].rend());
        Split(pts, ch[i], rev, b, maxLeaf, out);
        b += (long)rev.size();
    }
}

long FindBlock(const CCSRMat& A, long row, long col)
{
    if (row < 0 || row >= A.n || col < 0 || col >= A.n)
        return -1;
    for (int p = A.row_ptr[(size_t)row]; p < A.row_ptr[(size_t)row + 1]; ++p)
        if (A.col_index[(size_t)p] == col)
            return p;
    return -1;
}

}  // namespace

int InvMat(double* A, long n)
{
    std::vector<long> is((size_t)n + 1), js((size_t)n + 1);
    auto a = [&](long r, long c) -> double& { return A[(r - 1) * n + (c - 1)]; };
    for (long k = 1; k <= n; ++k) {
        double d = 0.0;
        for (long i = k; i <= n; ++i)
            for (long j = k; j <= n; ++j) {
                double t = std::sqrt(a(i, j) * a(i, j));
                if (t > d) {
                    d = t;
                    is[(size_t)k] = i;
                    js[(size_t)k] = j;
                }
            }
        if (d <= 0.0000000001)
            return -1;
        for (long j = 1; j <= n; ++j) {
            double t = a(k, j);
            a(k, j) = a(is[(size_t)k], j);
            a(is[(size_t)k], j) = t;
        }
        for (long i = 1; i <= n; ++i) {
            double t = a(i, k);
            a(i, k) = a(i, js[(size_t)k]);
            a(i, js[(size_t)k]) = t;
        }
        a(k, k) = 1.0 / a(k, k);
        for (long j = 1; j <= n; ++j)
            if (j != k)
                a(k, j) *= a(k, k);
        for (long i = 1; i <= n; ++i)
            if (i != k)
                for (long j = 1; j <= n; ++j)
                    if (j != k)
                        a(i, j) -= a(i, k) * a(k, j);
        for (long i = 1; i <= n; ++i)
            if (i != k)
                a(i, k) *= -1.0 * a(k, k);
    }
    for (long k = n; k >= 1; --k) {
        for (long j = 1; j <= n; ++j) {
            double t = a(k, j);
            a(k, j) = a(js[(size_t)k], j);
            a(js[(size_t)k], j) = t;
        }
        for (long i = 1; i <= n; ++i) {
            double t = a(i, k);
            a(i, k) = a(i, is[(size_t)k]);
            a(i, is[(size_t)k]) = t;
        }
    }
    return 1;
}

bool BuildMappedLeafPreConditioner(const double* centres, long N, long maxLeaf, const CCSRMat& A,
                                   const GlobalDofMap& dof, LeafPreconditioner& pre, const std::vector<long>* groups)
{
    pre = LeafPreconditioner();
    if (N <= 0 || maxLeaf <= 0 || A.n != N)
        return false;
    std::vector<long> bounds = {0, N};
    if (groups && groups->size() >= 2)
        bounds = *groups;
    if (bounds.front() != 0 || bounds.back() != N)
        return false;
    pre.repid.assign((size_t)N, -1);
    for (size_t g = 0; g + 1 < bounds.size(); ++g) {
        const long p0 = bounds[g], p1 = bounds[g + 1];
        if (p1 <= p0)
            return false;
        double rmin[3], rmax[3];
        for (int j = 0; j < 3; ++j)
            rmin[j] = rmax[j] = centres[3 * p0 + j];
        for (long i = p0 + 1; i < p1; ++i)
            for (int j = 0; j < 3; ++j) {
                if (rmin[j] > centres[3 * i + j])
                    rmin[j] = centres[3 * i + j];
                if (rmax[j] < centres[3 * i + j])
                    rmax[j] = centres[3 * i + j];
            }
        Cube root;
        double ml = rmax[0] - rmin[0];
        for (int j = 0; j < 3; ++j) {
            if (ml < (rmax[j] - rmin[j]))
                ml = rmax[j] - rmin[j];
            root.c[j] = rmin[j] + 0.5 * (rmax[j] - rmin[j]);
        }
        root.len = 1.02 * ml;
        std::vector<long> all((size_t)(p1 - p0));
        for (long i = p0; i < p1; ++i)
            all[(size_t)(i - p0)] = i;
        Split(centres, root, all, p0, maxLeaf, pre);
    }
    pre.leafCount = (long)pre.leafSize.size();
    long total = 0;
    for (long l = 0; l < pre.leafCount; ++l) {
        pre.matOffset.push_back(total);
        long n3 = 3 * pre.leafSize[(size_t)l];
        total += n3 * n3;
    }
    pre.inv.assign((size_t)total, 0.0);
    for (long l = 0; l < pre.leafCount; ++l) {
        long cnt = pre.leafSize[(size_t)l], n3 = 3 * cnt, b = pre.leafBegin[(size_t)l];
        double* P = &pre.inv[(size_t)pre.matOffset[(size_t)l]];
        for (long r = 0; r < cnt; ++r) {
            long row = pre.repid[(size_t)(b + r)];
            for (long c = 0; c < cnt; ++c) {
                long col = dof.GetPreferredUnknownBlockForRow(pre.repid[(size_t)(b + c)]);
                long p = FindBlock(A, row, col);
                if (p < 0)
                    continue;
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        P[(3 * r + i) * n3 + 3 * c + j] = A.value[(size_t)p * 9 + (size_t)(3 * i + j)];
            }
        }
        InvMat(P, n3);
    }
    return true;
}

}  // namespace tdbem
