// Lattice constants of the 18-velocity Eggels-Somers model shared by host and device code.
//
// Slot numbering, direction vectors and opposite pairs follow
// tdlbcpp/src/Header.h:35 (18 stored populations, no rest population) and
// tdlbcpp/src/Header.h:70-89 (Q01..Q18 and their (x,y,z) vectors).
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define TDLB_HD __host__ __device__
#else
#define TDLB_HD
#endif

namespace tdlb {

constexpr int kQ = 18;  // stored populations per cell

// Slot indices, 0-based, in the reference's order (Header.h:70-89).
enum Slot : int {
    Q01 = 0,  // +x
    Q02 = 1,  // -x
    Q03 = 2,  // +y
    Q04 = 3,  // -y
    Q05 = 4,  // +z
    Q06 = 5,  // -z
    Q07 = 6,  // +x +y
    Q08 = 7,  // -x -y
    Q09 = 8,  // +x +z
    Q10 = 9,  // -x -z
    Q11 = 10, // +y +z
    Q12 = 11, // -y -z
    Q13 = 12, // +x -y
    Q14 = 13, // -x +y
    Q15 = 14, // +x -z
    Q16 = 15, // -x +z
    Q17 = 16, // +y -z
    Q18 = 17, // -y +z
};

// Moment slots share the array with populations (Header.h:103-121).
enum MomentSlot : int { M01 = 0, M02, M03, M04, M05, M06, M07, M08, M09, M10,
                        M11, M12, M13, M14, M15, M16, M17, M18 };

// Opposite directions are adjacent pairs: Q01/Q02, Q03/Q04, ...
TDLB_HD constexpr int opposite(int d) { return d ^ 1; }

// Direction components per slot.
TDLB_HD constexpr int dir_x(int d) {
    return (d == Q01 || d == Q07 || d == Q09 || d == Q13 || d == Q15) ? 1
         : (d == Q02 || d == Q08 || d == Q10 || d == Q14 || d == Q16) ? -1 : 0;
}
TDLB_HD constexpr int dir_y(int d) {
    return (d == Q03 || d == Q07 || d == Q11 || d == Q14 || d == Q17) ? 1
         : (d == Q04 || d == Q08 || d == Q12 || d == Q13 || d == Q18) ? -1 : 0;
}
TDLB_HD constexpr int dir_z(int d) {
    return (d == Q05 || d == Q09 || d == Q11 || d == Q16 || d == Q18) ? 1
         : (d == Q06 || d == Q10 || d == Q12 || d == Q15 || d == Q17) ? -1 : 0;
}

// Esoteric-twist home offset: a population moving in direction d is stored at
// the cell shifted by +1 along every axis where d has a negative component
// (tdlbcpp/src/StreamingEsoteric.hpp:27-180 spells this out slot by slot).
TDLB_HD constexpr int twist_dx(int d) { return dir_x(d) < 0 ? 1 : 0; }
TDLB_HD constexpr int twist_dy(int d) { return dir_y(d) < 0 ? 1 : 0; }
TDLB_HD constexpr int twist_dz(int d) { return dir_z(d) < 0 ? 1 : 0; }

enum Layout : int32_t { LayoutIJKL = 0, LayoutLIJK = 1 };
enum Streaming : int32_t { StreamSimple = 0, StreamEsotwist = 1 };

}  // namespace tdlb
