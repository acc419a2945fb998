// hydra_faces.h -- the face list of the FP32-bulk HLLC-Batten form (host
// side, no device code): every face's flux is computed once and given to
// both of its cells.
#ifndef HYDRA_FACES_H
#define HYDRA_FACES_H

#include <cstddef>
#include <vector>

namespace hydra {

// nc: 4 x ne neighbour ids, face-major (face f of cell i at f*ne + i),
// 0-based, ghosts >= ne.  A cell lists its boundary faces and its faces
// towards higher-numbered cells, cell by cell: flist[k] = 4*cell + face,
// cell i's entries from fbase[i] to fbase[i+1] (ne+1 entries).  fslot[f*ne
// + i] = (k << 1) | s: cell i takes entry k for its face f, negated when s
// is 1 (the entry its neighbour listed).  A pair of cells that does not
// share exactly one face (not a valid mesh), or a cell that names itself,
// keeps one entry per side.
inline void build_face_tables(int ne, const int *nc, std::vector<int> &fslot, std::vector<int> &fbase,
                              std::vector<int> &flist)
{
    fslot.assign(4 * (size_t)ne, 0);
    fbase.assign((size_t)ne + 1, 0);
    flist.clear();
    flist.reserve(3 * (size_t)ne);
    // the face of j towards i when i and j share exactly one face, else -1
    auto back = [&](int i, int j) -> int {
        int nij = 0, nji = 0, fb = -1;
        for (int k = 0; k < 4; k++) {
            if (nc[(size_t)k * ne + i] == j) nij++;
            if (nc[(size_t)k * ne + j] == i) {
                nji++;
                fb = k;
            }
        }
        return (nij == 1 && nji == 1) ? fb : -1;
    };
    auto listed = [&](int i, int j) { return j >= ne || i <= j || back(i, j) < 0; };
    for (int i = 0; i < ne; i++) {
        fbase[i] = (int)flist.size();
        for (int f = 0; f < 4; f++) {
            if (listed(i, nc[(size_t)f * ne + i])) {
                fslot[(size_t)f * ne + i] = (int)flist.size() << 1;
                flist.push_back(4 * i + f);
            }
        }
    }
    fbase[ne] = (int)flist.size();
    for (int i = 0; i < ne; i++)
        for (int f = 0; f < 4; f++) {
            const int j = nc[(size_t)f * ne + i];
            if (listed(i, j)) continue;
            fslot[(size_t)f * ne + i] = fslot[(size_t)back(i, j) * ne + j] | 1;
        }
}

}  // namespace hydra

#endif
