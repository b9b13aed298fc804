// geneva_launchers.hpp -- header-only C++ shim with the reference's launcher
// names and signatures (include/miningPattern.h:1-31, include/reduction.h:1-4),
// forwarding to the C ABI of geneva_launchers.h.
//
// A driver written against the reference's headers (src/gpumine.cpp,
// src/gpumine_singleV.cpp) includes this header instead of miningPattern.h and
// reduction.h and links libgeneva_hip.so; its launcher calls stay as they are.
// The launches go to the context set with geneva_launchers::set_stream (one
// per host thread).  Like the reference launchers they return void; the first
// failing status on the thread is kept, as a sticky CUDA error is
// (src/gpumine.cpp:1066-1067), and geneva_launchers::take_error() returns and
// clears it.
#ifndef GENEVA_LAUNCHERS_HPP
#define GENEVA_LAUNCHERS_HPP

#include <sys/types.h>

#include "geneva_launchers.h"

namespace geneva_launchers {
inline gv_ctx_t& stream() {
    static thread_local gv_ctx_t s = nullptr;
    return s;
}
inline int& sticky() {
    static thread_local int e = GV_OK;
    return e;
}
inline void set_stream(gv_ctx_t ctx) { stream() = ctx; }
inline void note(int rc) {
    if (rc != GV_OK && sticky() == GV_OK) sticky() = rc;
}
inline int take_error() {
    const int e = sticky();
    sticky() = GV_OK;
    return e;
}
}  // namespace geneva_launchers

inline void initEmb_3V_1and2_NoHash(uint* vLabel_dev, uint* totWriteRowNum_dev, uint* edgeLabelPartition_dev,
                                    uint* neighborsData_dev, uint* recordPos1_dev, uint* recordPos2_dev,
                                    uint* newEmb_dev, uint intervalNum, uint totSrcNum, uint svidlabel,
                                    uint evid1label, uint evid2label, bool isExt1Restrict, bool isExt2Restrict,
                                    bool isExt1and2Restrict, bool isExt1and2SameLabel, bool isRecord1,
                                    bool isRecord2, bool isContinue, uint maxRowNum) {
    geneva_launchers::note(gv_initEmb_3V_1and2_NoHash(
        geneva_launchers::stream(), vLabel_dev, totWriteRowNum_dev, edgeLabelPartition_dev, neighborsData_dev,
        recordPos1_dev, recordPos2_dev, newEmb_dev, intervalNum, totSrcNum, svidlabel, evid1label, evid2label,
        isExt1Restrict, isExt2Restrict, isExt1and2Restrict, isExt1and2SameLabel, isRecord1, isRecord2,
        isContinue, maxRowNum));
}

inline void initEmb_2V_NoHash(uint* vLabel_dev, uint* totWriteRowNum_dev, uint* edgeLabelPartition_dev,
                              uint* neighborsData_dev, uint* recordPos_dev, uint* newEmb_dev, uint intervalNum,
                              uint totSrcNum, uint svidlabel, uint evidlabel, uint isRestrict, uint isRecord,
                              bool isContinue, uint maxRowNum) {
    geneva_launchers::note(gv_initEmb_2V_NoHash(geneva_launchers::stream(), vLabel_dev, totWriteRowNum_dev,
                                                edgeLabelPartition_dev, neighborsData_dev, recordPos_dev,
                                                newEmb_dev, intervalNum, totSrcNum, svidlabel, evidlabel,
                                                isRestrict, isRecord, isContinue, maxRowNum));
}

inline void extEmb_2V_2src_1and2_sameLabel_NoHash(uint* vLabel_dev, uint* totWriteRowNum_dev,
                                                  uint* edgeLabelPartition_dev, uint* neighborsData_dev,
                                                  uint* recordPos_dev, uint* auxArray_dev, uint* newEmb_dev,
                                                  uint* partialEmb_dev, uint intervalNum, uint partialRowNum,
                                                  uint embLen, bool isExt1Restrict, bool isExt2Restrict,
                                                  bool isExt1and2Restrict, bool isRecord, bool useRecord,
                                                  bool isContinue, uint maxRowNum) {
    geneva_launchers::note(gv_extEmb_2V_2src_1and2_sameLabel_NoHash(
        geneva_launchers::stream(), vLabel_dev, totWriteRowNum_dev, edgeLabelPartition_dev, neighborsData_dev,
        recordPos_dev, auxArray_dev, newEmb_dev, partialEmb_dev, intervalNum, partialRowNum, embLen,
        isExt1Restrict, isExt2Restrict, isExt1and2Restrict, isRecord, useRecord, isContinue, maxRowNum));
}

inline void extEmb_2V_2src_1and2_notSameLabel_NoHash(uint* vLabel_dev, uint* totWriteRowNum_dev,
                                                     uint* edgeLabelPartition_dev, uint* neighborsData_dev,
                                                     uint* recordPos1_dev, uint* recordPos2_dev, uint* auxArray_dev,
                                                     uint* newEmb_dev, uint* partialEmb_dev, uint intervalNum,
                                                     uint partialRowNum, uint embLen, bool isExt1Restrict,
                                                     bool isExt2Restrict, bool isRecord1, bool isRecord2,
                                                     bool useRecord1, bool useRecord2, bool isContinue,
                                                     uint maxRowNum) {
    geneva_launchers::note(gv_extEmb_2V_2src_1and2_notSameLabel_NoHash(
        geneva_launchers::stream(), vLabel_dev, totWriteRowNum_dev, edgeLabelPartition_dev, neighborsData_dev,
        recordPos1_dev, recordPos2_dev, auxArray_dev, newEmb_dev, partialEmb_dev, intervalNum, partialRowNum,
        embLen, isExt1Restrict, isExt2Restrict, isRecord1, isRecord2, useRecord1, useRecord2, isContinue,
        maxRowNum));
}

inline void extEmb_2V_1src_1and2_sameLabel_NoHash(uint* vLabel_dev, uint* totWriteRowNum_dev,
                                                  uint* edgeLabelPartition_dev, uint* neighborsData_dev,
                                                  uint* recordPos_dev, uint* auxArray_dev, uint* newEmb_dev,
                                                  uint* partialEmb_dev, uint intervalNum, uint partialRowNum,
                                                  uint embLen, bool isExt1Restrict, bool isExt2Restrict,
                                                  bool isExt1and2Restrict, bool isRecord, bool useRecord,
                                                  bool isContinue, uint maxRowNum) {
    geneva_launchers::note(gv_extEmb_2V_1src_1and2_sameLabel_NoHash(
        geneva_launchers::stream(), vLabel_dev, totWriteRowNum_dev, edgeLabelPartition_dev, neighborsData_dev,
        recordPos_dev, auxArray_dev, newEmb_dev, partialEmb_dev, intervalNum, partialRowNum, embLen,
        isExt1Restrict, isExt2Restrict, isExt1and2Restrict, isRecord, useRecord, isContinue, maxRowNum));
}

inline void extEmb_2V_1src_1and2_notSameLabel_NoHash(uint* vLabel_dev, uint* totWriteRowNum_dev,
                                                     uint* edgeLabelPartition_dev, uint* neighborsData_dev,
                                                     uint* recordPos1_dev, uint* recordPos2_dev, uint* auxArray_dev,
                                                     uint* newEmb_dev, uint* partialEmb_dev, uint intervalNum,
                                                     uint partialRowNum, uint embLen, bool isExt1Restrict,
                                                     bool isExt2Restrict, bool isRecord1, bool isRecord2,
                                                     bool useRecord1, bool useRecord2, bool isContinue,
                                                     uint maxRowNum) {
    geneva_launchers::note(gv_extEmb_2V_1src_1and2_notSameLabel_NoHash(
        geneva_launchers::stream(), vLabel_dev, totWriteRowNum_dev, edgeLabelPartition_dev, neighborsData_dev,
        recordPos1_dev, recordPos2_dev, auxArray_dev, newEmb_dev, partialEmb_dev, intervalNum, partialRowNum,
        embLen, isExt1Restrict, isExt2Restrict, isRecord1, isRecord2, useRecord1, useRecord2, isContinue,
        maxRowNum));
}

inline void extEmb_1V_NoHash(uint* vLabel_dev, uint* totWriteRowNum_dev, uint* edgeLabelPartition_dev,
                             uint* neighborsData_dev, uint* recordPos_dev, uint* auxArray_dev, uint* newEmb_dev,
                             uint* partialEmb_dev, uint intervalNum, uint partialRowNum, uint embLen,
                             bool isExtRestrict, bool isRecord, bool useRecord, bool isLastPhase, bool isContinue,
                             uint maxRowNum) {
    geneva_launchers::note(gv_extEmb_1V_NoHash(geneva_launchers::stream(), vLabel_dev, totWriteRowNum_dev,
                                               edgeLabelPartition_dev, neighborsData_dev, recordPos_dev,
                                               auxArray_dev, newEmb_dev, partialEmb_dev, intervalNum,
                                               partialRowNum, embLen, isExtRestrict, isRecord, useRecord,
                                               isLastPhase, isContinue, maxRowNum));
}

inline void reductionPhase(uint* vLabel, uint* totWriteRowNum, uint* edgeLabelPartition, uint* neighborsData,
                           uint* baseRecordPos, uint* auxArray, uint* partialEmb, uint intervalNum,
                           uint partialRowNum, uint embLen, bool isLastPhase) {
    geneva_launchers::note(gv_reductionPhase(geneva_launchers::stream(), vLabel, totWriteRowNum,
                                             edgeLabelPartition, neighborsData, baseRecordPos, auxArray,
                                             partialEmb, intervalNum, partialRowNum, embLen, isLastPhase));
}

inline void copyPartialEmb(uint* src_dev, uint* dst_dev, uint numUints) {
    geneva_launchers::note(gv_copyPartialEmb(geneva_launchers::stream(), src_dev, dst_dev, numUints));
}

#endif  // GENEVA_LAUNCHERS_HPP
