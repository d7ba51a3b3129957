// rd_jni.cpp -- JNI entry point of libCReaDDyGPU.so (lib/ReaDDyGPU.cpp:31-503).
//
// Same protocol as the reference: read the six input files, then for every
// frame run stepsPerFrame BD steps, hand the positions to
// TopGPU.frameCallback(int frame, float[] JPos) with
// JPos = [N, (x, y, z, index, readdyId) x N] (lib/ReaDDyGPU.cpp:264-292), and
// apply the reactions Java leaves in the field float[] cReactions =
// [count, (readdyId, newType, x, y, z, index) x count] (lib/ReaDDyGPU.cpp:
// 303-463).  Errors are printed and end the call, as in the reference.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/rdgpu.h"
#include "../../include/readdy_jni.h"

namespace {

// messages go to stdout, flushed at once as the reference's cout << endl
void say(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void say(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vprintf(fmt, ap);
    va_end(ap);
    fflush(stdout);
}

typedef jclass (*GetObjectClass_t)(JNIEnv*, jobject);
typedef jmethodID (*GetMethodID_t)(JNIEnv*, jclass, const char*, const char*);
typedef jboolean (*CallBooleanMethod_t)(JNIEnv*, jobject, jmethodID, ...);
typedef jfieldID (*GetFieldID_t)(JNIEnv*, jclass, const char*, const char*);
typedef jobject (*GetObjectField_t)(JNIEnv*, jobject, jfieldID);
typedef const char* (*GetStringUTFChars_t)(JNIEnv*, jstring, jboolean*);
typedef void (*ReleaseStringUTFChars_t)(JNIEnv*, jstring, const char*);
typedef jsize (*GetArrayLength_t)(JNIEnv*, jarray);
typedef jfloatArray (*NewFloatArray_t)(JNIEnv*, jsize);
typedef void* (*GetPrimitiveArrayCritical_t)(JNIEnv*, jarray, jboolean*);
typedef void (*ReleasePrimitiveArrayCritical_t)(JNIEnv*, jarray, void*, jint);
typedef void (*DeleteLocalRef_t)(JNIEnv*, jobject);

template <class T>
T fn(JNIEnv* env, int idx)
{
    return reinterpret_cast<T>(const_cast<void*>((*env)[idx]));
}

bool jstr(JNIEnv* env, jstring s, std::string& out)
{
    if (!s) return false;
    const char* c = fn<GetStringUTFChars_t>(env, RD_JNI_GetStringUTFChars)(env, s, nullptr);
    if (!c) return false;
    out = c;
    fn<ReleaseStringUTFChars_t>(env, RD_JNI_ReleaseStringUTFChars)(env, s, c);
    return true;
}

}  // namespace

extern "C" void Java_readdy_impl_sim_top_TopGPU_cCreateSimulation(JNIEnv* env, jobject obj, jboolean testmode,
                                                                 jstring paramParticles, jstring tplgyPotentials,
                                                                 jstring paramGroups, jstring tplgyCoord,
                                                                 jstring tplgyGroups, jstring paramGlobal)
{
    (void)testmode;
    std::string p_part, p_pot, p_grp, p_crd, p_tgrp, p_glob;
    if (!jstr(env, paramParticles, p_part) || !jstr(env, tplgyPotentials, p_pot) || !jstr(env, paramGroups, p_grp) ||
        !jstr(env, tplgyCoord, p_crd) || !jstr(env, tplgyGroups, p_tgrp) || !jstr(env, paramGlobal, p_glob)) {
        say("error while reading the input file paths\n");
        return;
    }
    rd_model* model = nullptr;
    char err[512] = { 0 };
    if (rd_model_read_xml(p_part.c_str(), p_pot.c_str(), p_grp.c_str(), p_crd.c_str(), p_tgrp.c_str(),
                          p_glob.c_str(), &model, err, (int)sizeof(err)) != RD_OK) {
        say("error while reading the input files: %s\n", err);
        return;
    }
    rd_model_info info;
    rd_model_get_info(model, &info);
    if (info.n_groups > 0)
        say("%d particle groups: %d bonded pairs\n", info.n_groups, info.n_bonds);

    // lib/ReaDDyGPU.cpp:101: maxCutoff = 10; the reference seeds its RNG with
    // the wall clock (CudaSimulation.cu:773), RDGPU_SEED fixes the seed.
    const char* seed_env = getenv("RDGPU_SEED");
    uint64_t seed = seed_env ? strtoull(seed_env, nullptr, 10) : (uint64_t)time(nullptr);
    rd_system* h = nullptr;
    if (rd_create_from_model(model, 10.0, seed, &h) != RD_OK) {
        say("error during initialization\n");
        rd_model_free(model);
        return;
    }
    std::vector<int> rid(info.n_particles);
    rd_model_particles(model, nullptr, nullptr, rid.data());
    rd_model_free(model);

    jclass cls = fn<GetObjectClass_t>(env, RD_JNI_GetObjectClass)(env, obj);
    jmethodID mid = cls ? fn<GetMethodID_t>(env, RD_JNI_GetMethodID)(env, cls, "frameCallback", "(I[F)Z") : nullptr;
    if (!mid) {
        say("mthod not found\n");
        rd_destroy(h);
        return;
    }
    jfieldID fid = fn<GetFieldID_t>(env, RD_JNI_GetFieldID)(env, cls, "cReactions", "[F");
    say("%d frames with each %d steps\n", info.frames, info.steps_per_frame);

    // Frames are handed off without stopping the GPU: the positions of frame f
    // are copied (pinned, asynchronous) while the steps of frame f+1 already
    // run.  If the Java side then returns reactions for frame f, those steps
    // are discarded (rd_frame_rewind) and rerun from the changed state, with
    // the same noise.  Speculation is used while the previous frame had no
    // reactions, so reactive runs do not pay for discarded frames.
    // Trajectory files: the reference's JNI entry has them switched off
    // (output = false, lib/ReaDDyGPU.cpp:123) and its stand-alone driver writes
    // ReaDDy_output/pdb.pdb and out_traj.xml per frame; RDGPU_OUTPUT=<dir>
    // turns them on here, written while the next frame runs.
    rd_traj* pdb = nullptr;
    rd_traj* xmlt = nullptr;
    if (const char* od = getenv("RDGPU_OUTPUT")) {
        std::string d(od);
        if (rd_traj_open((d + "/pdb.pdb").c_str(), RD_TRAJ_PDB, &pdb) != RD_OK ||
            rd_traj_open((d + "/out_traj.xml").c_str(), RD_TRAJ_XML, &xmlt) != RD_OK) {
            say("error while opening trajectory output files\n");
            if (pdb) rd_traj_close(pdb);
            rd_destroy(h);
            return;
        }
    }
    const int spf = info.steps_per_frame;
    bool speculate = true;
    if (info.frames > 0 && rd_simulate_async(h, spf) != RD_OK) {
        say("error: %s\n", rd_last_error(h));
        rd_destroy(h);
        return;
    }
    for (int frame = 0; frame < info.frames; ++frame) {
        if (rd_frame_copy(h) != RD_OK) {
            say("error while copying positions from device: %s\n", rd_last_error(h));
            break;
        }
        const bool ahead = speculate && frame + 1 < info.frames;
        if (ahead && rd_simulate_async(h, spf) != RD_OK) {
            say("error: %s\n", rd_last_error(h));
            break;
        }
        const float* xyz = nullptr;
        const int* type = nullptr;
        int n = 0;
        if (rd_frame_wait(h, &xyz, &type, &n) != RD_OK) {
            say("error: %s\n", rd_last_error(h));
            break;
        }
        say("frame: %d (step %lld)\n", frame, (long long)frame * spf);
        if (pdb && (rd_traj_write_frame(pdb, frame, n, xyz, type) != RD_OK ||
                    rd_traj_write_frame(xmlt, frame, n, xyz, type) != RD_OK)) {
            say(" error while writing trajectory output\n");
            break;
        }
        // JPos = [N, (x, y, z, index, readdyId) x N]
        jfloatArray jpos = fn<NewFloatArray_t>(env, RD_JNI_NewFloatArray)(env, (jsize)(5 * n + 1));
        float* pos = jpos ? (float*)fn<GetPrimitiveArrayCritical_t>(env, RD_JNI_GetPrimitiveArrayCritical)(env, jpos, nullptr)
                          : nullptr;
        if (!pos) {
            say("cant get array pointer\n");
            break;
        }
        pos[0] = (float)n;
        for (int j = 0; j < n; ++j) {
            pos[1 + 5 * j] = xyz[3 * j];
            pos[2 + 5 * j] = xyz[3 * j + 1];
            pos[3 + 5 * j] = xyz[3 * j + 2];
            pos[4 + 5 * j] = (float)j;
            pos[5 + 5 * j] = (float)rid[j];
        }
        fn<ReleasePrimitiveArrayCritical_t>(env, RD_JNI_ReleasePrimitiveArrayCritical)(env, jpos, pos, 0);
        jboolean ok = fn<CallBooleanMethod_t>(env, RD_JNI_CallBooleanMethod)(env, obj, mid, (jint)frame, jpos);
        fn<DeleteLocalRef_t>(env, RD_JNI_DeleteLocalRef)(env, jpos);
        if (!ok) {
            say("error in java\n");
            break;
        }
        // reactions from the Java side
        std::vector<float> cr;
        if (fid) {
            jobject arr = fn<GetObjectField_t>(env, RD_JNI_GetObjectField)(env, obj, fid);
            if (arr) {
                jsize len = fn<GetArrayLength_t>(env, RD_JNI_GetArrayLength)(env, arr);
                float* r = (float*)fn<GetPrimitiveArrayCritical_t>(env, RD_JNI_GetPrimitiveArrayCritical)(env, arr, nullptr);
                if (r && len > 0) {
                    int cnt = (int)r[0];
                    if (cnt > 0 && 1 + 6 * (long)cnt <= len) cr.assign(r, r + 1 + 6 * (size_t)cnt);
                    else if (cnt > 0) say("error: cReactions shorter than its count\n");
                }
                if (r) fn<ReleasePrimitiveArrayCritical_t>(env, RD_JNI_ReleasePrimitiveArrayCritical)(env, arr, r, 0);
                fn<DeleteLocalRef_t>(env, RD_JNI_DeleteLocalRef)(env, arr);
            }
        }
        if (cr.empty()) {
            speculate = true;
            if (!ahead && frame + 1 < info.frames && rd_simulate_async(h, spf) != RD_OK) {
                say("error: %s\n", rd_last_error(h));
                break;
            }
            continue;
        }
        speculate = false;
        if (ahead && rd_frame_rewind(h) != RD_OK) {
            say("error: %s\n", rd_last_error(h));
            break;
        }
        // same compaction as rd_apply_reactions, applied to the ReaDDy ids
        std::vector<int> keep(rid.begin(), rid.end());
        std::vector<char> dead(n, 0);
        const int cnt = (int)cr[0];
        std::vector<int> born;
        for (int k = 0; k < cnt; ++k) {
            const float* e = &cr[1 + 6 * k];
            int index = (int)e[5], newType = (int)e[1];
            if (index >= 0 && index < n) dead[index] = 1;
            if (newType != -1) born.push_back((int)e[0]);
        }
        if (rd_apply_reactions(h, cr.data()) != RD_OK) {
            say("error: %s\n", rd_last_error(h));
            break;
        }
        rid.clear();
        for (int j = 0; j < n; ++j)
            if (!dead[j]) rid.push_back(keep[j]);
        rid.insert(rid.end(), born.begin(), born.end());
        if (frame + 1 < info.frames && rd_simulate_async(h, spf) != RD_OK) {
            say("error: %s\n", rd_last_error(h));
            break;
        }
    }
    if (pdb) rd_traj_close(pdb);
    if (xmlt) rd_traj_close(xmlt);
    rd_destroy(h);
}
