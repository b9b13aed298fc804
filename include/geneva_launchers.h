/*
 * geneva_launchers.h -- the reference's kernel launchers as a C ABI
 * (libgeneva_hip.so).
 *
 * One entry per launcher of include/miningPattern.h:1-31 and
 * include/reduction.h:1-4, with the launcher's parameter list unchanged after
 * a leading context (the stream the launch is ordered on), returning a status
 * (GV_OK, or a GV_ERR_* of geneva_hip.h) instead of void.  They work on the
 * reference's device layout: vLabel_dev[v] = label of vertex v; a partition
 * block as the reference driver uploads it (edgeLabelPartition_dev = its
 * array0, neighborsData_dev = its array3, skip-index prefixes in place:
 * tools/genDatabase.cpp:437-580); auxArray_dev as src/gpumine.cpp:40-61 lays
 * it out, spanning 64 uints as the reference kernels read it
 * (src/miningPattern.cu:2044-2046); row-major uint32 tables in which a row
 * holding 0 is invalid and skipped; totWriteRowNum_dev[0] = rows written
 * (rows are placed from that index on), [1] = rows not written after an
 * overflow.  The caller owns every device buffer, as in the reference; the
 * context keeps a workspace of a few words per input row (grown on demand,
 * or sized once by gv_launchers_reserve).
 *
 * Differences from the reference launchers (the rows each call leaves are
 * the valid rows the reference leaves, in another order):
 *  - a candidate equal to a mapped same-label vertex is not written, where the
 *    reference writes the row with a 0 (src/miningPattern.cu:190-194), so
 *    totWriteRowNum_dev[0] counts valid rows only, in every phase;
 *  - a launch that would pass row maxRowNum - 1 stops there, raises
 *    totWriteRowNum_dev[0] to at least maxRowNum and sets [1] so that
 *    [0] - [1] = rows written (the reference's overflow protocol,
 *    src/gpumine_singleV.cpp:258-279); the same call with isContinue = true
 *    resumes from the context's resume point, not from records behind the
 *    table;
 *  - recordPos caches are written when isRecord* is set and never read
 *    (useRecord* has no effect);
 *  - reductionPhase compacts survivors to the front of partialEmb in every
 *    phase (the reference leaves 0-marked holes when !isLastPhase,
 *    src/reduction.cu:230-262), and decodes pairs past the third correctly
 *    (the reference shifts by the word index, src/reduction.cu:201-202).
 *
 * Calls on one context must not overlap.  Each call synchronises the
 * context's stream once or twice (to read sizes), as the reference driver's
 * read-back after every launch does.
 */
#ifndef GENEVA_LAUNCHERS_H
#define GENEVA_LAUNCHERS_H

#include <stdbool.h>
#include <stdint.h>

#include "geneva_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* initEmb_2V_NoHash (include/miningPattern.h:6-7; src/miningPattern.cu:3417-3426):
 * rows (s, e) for every source index of the partition whose vertex s carries
 * svidlabel and every e of label evidlabel in its list, e < s if isRestrict. */
int gv_initEmb_2V_NoHash(gv_ctx_t ctx, uint32_t* vLabel_dev, uint32_t* totWriteRowNum_dev,
                         uint32_t* edgeLabelPartition_dev, uint32_t* neighborsData_dev,
                         uint32_t* recordPos_dev, uint32_t* newEmb_dev, uint32_t intervalNum,
                         uint32_t totSrcNum, uint32_t svidlabel, uint32_t evidlabel, uint32_t isRestrict,
                         uint32_t isRecord, bool isContinue, uint32_t maxRowNum);

/* initEmb_3V_1and2_NoHash (include/miningPattern.h:1-4): rows (s, e1, e2),
 * e1 of label evid1label and e2 of label evid2label from s's list; e1 < s,
 * e2 < s, e2 < e1 (same label only) as flagged; e1 != e2. */
int gv_initEmb_3V_1and2_NoHash(gv_ctx_t ctx, uint32_t* vLabel_dev, uint32_t* totWriteRowNum_dev,
                               uint32_t* edgeLabelPartition_dev, uint32_t* neighborsData_dev,
                               uint32_t* recordPos1_dev, uint32_t* recordPos2_dev, uint32_t* newEmb_dev,
                               uint32_t intervalNum, uint32_t totSrcNum, uint32_t svidlabel,
                               uint32_t evid1label, uint32_t evid2label, bool isExt1Restrict,
                               bool isExt2Restrict, bool isExt1and2Restrict, bool isExt1and2SameLabel,
                               bool isRecord1, bool isRecord2, bool isContinue, uint32_t maxRowNum);

/* extEmb_2V_2src_1and2_sameLabel_NoHash (include/miningPattern.h:9-12) and
 * the three other two-vertex extensions (:14-27): each input row of
 * embLen - 2 columns gets e1 (label auxArray[0]) from row[auxArray[4]]'s list
 * and e2 (label auxArray[1]) from row[auxArray[5]]'s list; e1 < row[aux[2]]
 * and e2 < row[aux[3]] as flagged, e2 < e1 (same label only) as flagged,
 * e1 != e2, and each differs from the mapped same-label columns aux[8..]. */
int gv_extEmb_2V_2src_1and2_sameLabel_NoHash(gv_ctx_t ctx, uint32_t* vLabel_dev, uint32_t* totWriteRowNum_dev,
                                             uint32_t* edgeLabelPartition_dev, uint32_t* neighborsData_dev,
                                             uint32_t* recordPos_dev, uint32_t* auxArray_dev, uint32_t* newEmb_dev,
                                             uint32_t* partialEmb_dev, uint32_t intervalNum,
                                             uint32_t partialRowNum, uint32_t embLen, bool isExt1Restrict,
                                             bool isExt2Restrict, bool isExt1and2Restrict, bool isRecord,
                                             bool useRecord, bool isContinue, uint32_t maxRowNum);
int gv_extEmb_2V_2src_1and2_notSameLabel_NoHash(gv_ctx_t ctx, uint32_t* vLabel_dev, uint32_t* totWriteRowNum_dev,
                                                uint32_t* edgeLabelPartition_dev, uint32_t* neighborsData_dev,
                                                uint32_t* recordPos1_dev, uint32_t* recordPos2_dev,
                                                uint32_t* auxArray_dev, uint32_t* newEmb_dev,
                                                uint32_t* partialEmb_dev, uint32_t intervalNum,
                                                uint32_t partialRowNum, uint32_t embLen, bool isExt1Restrict,
                                                bool isExt2Restrict, bool isRecord1, bool isRecord2,
                                                bool useRecord1, bool useRecord2, bool isContinue,
                                                uint32_t maxRowNum);
int gv_extEmb_2V_1src_1and2_sameLabel_NoHash(gv_ctx_t ctx, uint32_t* vLabel_dev, uint32_t* totWriteRowNum_dev,
                                             uint32_t* edgeLabelPartition_dev, uint32_t* neighborsData_dev,
                                             uint32_t* recordPos_dev, uint32_t* auxArray_dev, uint32_t* newEmb_dev,
                                             uint32_t* partialEmb_dev, uint32_t intervalNum,
                                             uint32_t partialRowNum, uint32_t embLen, bool isExt1Restrict,
                                             bool isExt2Restrict, bool isExt1and2Restrict, bool isRecord,
                                             bool useRecord, bool isContinue, uint32_t maxRowNum);
int gv_extEmb_2V_1src_1and2_notSameLabel_NoHash(gv_ctx_t ctx, uint32_t* vLabel_dev, uint32_t* totWriteRowNum_dev,
                                                uint32_t* edgeLabelPartition_dev, uint32_t* neighborsData_dev,
                                                uint32_t* recordPos1_dev, uint32_t* recordPos2_dev,
                                                uint32_t* auxArray_dev, uint32_t* newEmb_dev,
                                                uint32_t* partialEmb_dev, uint32_t intervalNum,
                                                uint32_t partialRowNum, uint32_t embLen, bool isExt1Restrict,
                                                bool isExt2Restrict, bool isRecord1, bool isRecord2,
                                                bool useRecord1, bool useRecord2, bool isContinue,
                                                uint32_t maxRowNum);

/* extEmb_1V_NoHash (include/miningPattern.h:29-31): each input row of
 * embLen - 1 columns gets every e of label auxArray[0] in
 * row[auxArray[4]]'s list, e < row[auxArray[2]] if isExtRestrict, e different
 * from the mapped same-label columns auxArray[8 .. 8 + auxArray[6]). */
int gv_extEmb_1V_NoHash(gv_ctx_t ctx, uint32_t* vLabel_dev, uint32_t* totWriteRowNum_dev,
                        uint32_t* edgeLabelPartition_dev, uint32_t* neighborsData_dev, uint32_t* recordPos_dev,
                        uint32_t* auxArray_dev, uint32_t* newEmb_dev, uint32_t* partialEmb_dev,
                        uint32_t intervalNum, uint32_t partialRowNum, uint32_t embLen, bool isExtRestrict,
                        bool isRecord, bool useRecord, bool isLastPhase, bool isContinue,
                        uint32_t maxRowNum);

/* reductionPhase (include/reduction.h:1-2; src/reduction.cu:266-278): keeps
 * the rows of partialEmb (partialRowNum rows of embLen) in which every pair
 * (start column, end column) of auxArray is an edge of this partition, end
 * vertex in the start vertex's group of the pair's label; survivors move to
 * the front of partialEmb.  totWriteRowNum[0] += survivors when isLastPhase,
 * else = max(itself, survivors). */
int gv_reductionPhase(gv_ctx_t ctx, uint32_t* vLabel, uint32_t* totWriteRowNum, uint32_t* edgeLabelPartition,
                      uint32_t* neighborsData, uint32_t* baseRecordPos, uint32_t* auxArray,
                      uint32_t* partialEmb, uint32_t intervalNum, uint32_t partialRowNum, uint32_t embLen,
                      bool isLastPhase);

/* copyPartialEmb (include/reduction.h:4; src/reduction.cu:280-296): copy
 * numUints uints, overlapping ranges allowed in either direction. */
int gv_copyPartialEmb(gv_ctx_t ctx, uint32_t* src_dev, uint32_t* dst_dev, uint32_t numUints);

/* gv_extEmb_1V_NoHash for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, hosts that forbid allocation): no
 * counterpart in the reference.  Same arguments, results and overflow
 * protocol; the working arrays live in a fixed device array per GPU (about
 * 64 MB, taken once per process), so at most 2,000,000 input rows
 * (partialRowNum; more gives GV_ERR_CAPACITY), and the call allocates no
 * device, pinned or host-heap memory.  All calls on one GPU share that array
 * and one resume point: calls that overlap on one GPU race. */
int gv_extEmb_1V_NoHash_noalloc(gv_ctx_t ctx, uint32_t* vLabel_dev, uint32_t* totWriteRowNum_dev,
                                uint32_t* edgeLabelPartition_dev, uint32_t* neighborsData_dev,
                                uint32_t* recordPos_dev, uint32_t* auxArray_dev, uint32_t* newEmb_dev,
                                uint32_t* partialEmb_dev, uint32_t intervalNum, uint32_t partialRowNum,
                                uint32_t embLen, bool isExtRestrict, bool isRecord, bool useRecord,
                                bool isLastPhase, bool isContinue, uint32_t maxRowNum);

/* Size the context's launcher workspace for input tables of up to `rows`
 * rows and reductions of up to `reduction_words` uints, so later calls
 * allocate nothing.  No counterpart in the reference. */
int gv_launchers_reserve(gv_ctx_t ctx, uint64_t rows, uint64_t reduction_words);

#ifdef __cplusplus
}
#endif
#endif /* GENEVA_LAUNCHERS_H */
