/* Compares the two layout tables of layout_probe.c; exit 0 when equal. */
#include <stddef.h>
#include <stdio.h>

extern const size_t layout_reference[], layout_rover[];
extern const size_t layout_reference_count, layout_rover_count;

int main(void) {
    int bad = layout_reference_count != layout_rover_count;
    for (size_t i = 0; !bad && i < layout_rover_count; i++) {
        if (layout_reference[i] != layout_rover[i]) {
            printf("entry %zu: reference %zu, rover_gpu.h %zu\n", i, layout_reference[i],
                   layout_rover[i]);
            bad = 1;
        }
    }
    printf("%zu layout entries %s\n", layout_rover_count, bad ? "DIFFER" : "equal");
    return bad;
}
