/* Field component selector (drop-in for engine/fdtd/field_component.h:1-8). */
#ifndef RFSYN_FDTD_FIELD_COMPONENT_H
#define RFSYN_FDTD_FIELD_COMPONENT_H

typedef enum {
    FIELD_EX, FIELD_EY, FIELD_EZ,
    FIELD_HX, FIELD_HY, FIELD_HZ
} field_component_t;

#endif
