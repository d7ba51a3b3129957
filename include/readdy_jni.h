/*
 * readdy_jni.h -- the JNI entry point of libCReaDDyGPU.so.
 *
 * Replaces lib/simulationGPU.h:15-16 (javah output for
 * readdy.impl.sim.top.TopGPU.cCreateSimulation, JVM signature
 * (ZLjava/lang/String;Ljava/lang/String;Ljava/lang/String;Ljava/lang/String;
 *  Ljava/lang/String;Ljava/lang/String;)V), defined in lib/ReaDDyGPU.cpp:31.
 *
 * No JDK is installed here, so the few JNI types the entry point needs are
 * declared below.  A JNIEnv* points to a pointer to the JNI function table;
 * the entry point calls through the table by the indices that the Java
 * Native Interface Specification ("Interface Function Table") assigns.
 */
#ifndef READDY_JNI_H
#define READDY_JNI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t jint;
typedef int32_t jsize;
typedef uint8_t jboolean;
typedef float jfloat;
typedef void* jobject;
typedef jobject jclass;
typedef jobject jstring;
typedef jobject jarray;
typedef jarray jfloatArray;
typedef void* jmethodID;
typedef void* jfieldID;
typedef const void* const* JNIEnv;   /* *env is the function table */

/* function-table indices (JNI specification, Interface Function Table) */
enum {
    RD_JNI_DeleteLocalRef = 23,
    RD_JNI_GetObjectClass = 31,
    RD_JNI_GetMethodID = 33,
    RD_JNI_CallBooleanMethod = 37,
    RD_JNI_GetFieldID = 94,
    RD_JNI_GetObjectField = 95,
    RD_JNI_GetStringUTFChars = 169,
    RD_JNI_ReleaseStringUTFChars = 170,
    RD_JNI_GetArrayLength = 171,
    RD_JNI_NewFloatArray = 181,
    RD_JNI_GetPrimitiveArrayCritical = 222,
    RD_JNI_ReleasePrimitiveArrayCritical = 223,
    RD_JNI_ExceptionCheck = 228,
    RD_JNI_TABLE_SIZE = 233
};

/* lib/ReaDDyGPU.cpp:31: reads the six input files, runs all frames, calls
 * back frameCallback(int, float[]) after each frame and applies the
 * reactions found in the field float[] cReactions. */
__attribute__((visibility("default"))) void Java_readdy_impl_sim_top_TopGPU_cCreateSimulation(
    JNIEnv* env, jobject obj, jboolean testmode, jstring paramParticles, jstring tplgyPotentials,
    jstring paramGroups, jstring tplgyCoord, jstring tplgyGroups, jstring paramGlobal);

#ifdef __cplusplus
}
#endif
#endif
