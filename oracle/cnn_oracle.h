/* CPU oracle for the learned disparity-proposal network -- TEST INFRASTRUCTURE ONLY. */
#ifndef CLUTEC_CNN_ORACLE_H
#define CLUTEC_CNN_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif
int orc_cnn_version(void);
#ifdef __cplusplus
}
#endif
#endif
