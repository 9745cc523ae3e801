// Splitting one site off the optimised multisite tensor by a truncated
// per-sector SVD (src/siteTensor_decompose.c:587-1444, selectS 1196-1289).
// The host builds the sector matrices' layout and makes the truncation
// choice; the SVDs and the data movement run on the device
// (dev/decompose_dev.hip).
#pragma once
#include <cstdint>
#include <vector>

#include "model.hpp"
#include "sym.hpp"
#include "tensors.hpp"

namespace t3 {

struct SvalSelect {  // include/siteTensor.h:425-431
    int minD = 0, maxD = 0;
    double truncerr = 0;
    char truncType = 'E';
};

struct SplitInput {
    SymInfo sym;
    int nrsites = 0;
    int siteV = -1;                        // index (in the multisite) of the site split off
    Symsecs ss[STEPSPECS_MSITES][3];       // sectors of every site's bonds (bond ids set)
    std::vector<QN> qnumbers;              // multisite: nrblocks * nrsites
    std::vector<int64_t> beginblock;       // multisite: nrblocks + 1
    SvalSelect sel;
};

// One sector of the cut bond: the A blocks it gathers, the U and V blocks
// it scatters to (in the reference's order) and the matrix dims.
struct SplitSector {
    std::vector<int> idA;            // A blocks of this sector
    std::vector<int> idU;            // U blocks (rows)
    std::vector<int64_t> Mstart;     // row offsets per U block (cut dim 1)
    std::vector<int> idV;            // V blocks (columns)
    std::vector<int64_t> Nstart;     // column offsets per V block
    int64_t M = 0, N = 0, dimS = 0;  // dimS = min(M, N)
};

struct SplitPlan {
    int id_siteV = -1, id_bond = -1, id_csite = -1, id_cbond = -1;
    int nrSss = 0;  // sectors of the cut bond
    // U: the other sites; V: the split site (unique quantum numbers, sorted)
    int U_nrsites = 0;
    std::vector<QN> U_qn, V_qn;
    std::vector<SplitSector> sec;
    // per A block: sector, row offset within its sector, column offset,
    // dims (before V, V site, after V) as in get_dims (src/siteTensor_decompose.c:917-953)
    struct ABlock {
        int sector;
        int64_t row0, col0;
        int64_t d[3];
    };
    std::vector<ABlock> ablocks;
};

SplitPlan plan_split(const SplitInput& in);

// selectS (src/siteTensor_decompose.c:1196-1289): per sector the number of
// singular values kept; norm and entropy before [0] and after [1].
struct SelectRes {
    double norm[2] = {0, 0}, entropy[2] = {0, 0};
};
std::vector<int> select_singular_values(const std::vector<std::vector<double>>& sing,
                                        const std::vector<int>& mult, const SvalSelect& sel,
                                        SelectRes* res);

// qr_step (src/siteTensor_decompose.c:1587-1630): the one-site tensor A is
// QR-decomposed over its bond to the next orthogonality centre (qr 313-355,
// qrblocks 270-311: per sector of that bond, the blocks with that sector
// stacked row-wise in block order) and R multiplied into the centre
// (multiplyR 416-490).
struct QRInput {
    SymInfo sym;
    Symsecs ssA[3], ssC[3];  // bonds of A and of the centre (bond ids set)
    const SiteTensor* A = nullptr;
    const SiteTensor* C = nullptr;
};
struct QRSector {
    std::vector<int> blocks;         // A blocks with this sector, in block order
    std::vector<int64_t> rowstart;   // row offset of each block
    int64_t M = 0, N = 0, K = 0;     // rows, columns (the sector's dim), min(M, N)
};
struct QRPlan {
    int bondA = -1, bondC = -1;  // the shared bond's position in A and in the centre
    std::vector<QRSector> sec;
    // per A block: sector, dims before / at / after the bond
    struct ABlock {
        int sector;
        int64_t before, after;
    };
    std::vector<ABlock> ablocks;
};
QRPlan plan_qr(const QRInput& in);

}  // namespace t3
