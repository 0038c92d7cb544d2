/* Session C-ABI, version 1 (same contract as Niryukti's include/vantage/c_api.h:1-14
 * and src/c_api.cpp:16-46): the caller owns the handle and frees it with
 * vantage_close; returned strings are borrowed until the next call on the same
 * handle; calls on one handle are serialised (by the library); no C++
 * exception crosses the ABI; on failure NULL is returned and
 * vantage_last_error() holds the calling thread's message.
 *
 * Requests are JSON, exactly Niryukti's session protocol (src/session.hpp:13-73):
 *   {"action":"solve","device","method","time_limit","tol","threads","iterations","warm_start"}
 *   {"action":"update","rows":{name:{"lb","ub"}},"variables":{...},"objective":{name:value}}
 * Unknown keys are rejected.  This engine runs PDHG on the GPU; a request for
 * device "cpu" or "cuda" answers status UNSUPPORTED, and the default device
 * is "gpu". */
#ifndef VANTAGE_C_API_H
#define VANTAGE_C_API_H
#ifdef __cplusplus
extern "C" {
#endif
typedef struct vantage_session vantage_session;
unsigned vantage_abi_version(void);
vantage_session *vantage_open(const char *model_path);
void vantage_close(vantage_session *session);
const char *vantage_last_error(void);
const char *vantage_request(vantage_session *session, const char *request_json);
#ifdef __cplusplus
}
#endif
#endif
